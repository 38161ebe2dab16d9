"""The LDS-staged BPs take every tap from the rectangle they staged; that
rectangle is the bounding box of the block's projected corners widened by
one texel.  These CPU tests replay, in float32, the footprint arithmetic
of `bp_footprint` (csrc/proj3d.hip) and `bp2d_footprint`
(csrc/proj2d_dev.hpp) and the taps of the kernels (the same coefficient
layout, fused multiply-adds and incremental z / row steps) over whole
geometries, and check that every tap falls inside its staged rectangle --
the property the kernels' index clamp relies on to never change a value.
Planted mistakes (the footprint of the block's first z slice only, of the
tile's first row only) must be caught.
"""
import math

import numpy as np
import pytest

from oracle import geometry as OG, proj3d as O3

f32 = np.float32


def fma(a, b, c):
    """fp32 fused multiply-add (the product of two floats is exact in
    float64; one rounding of the sum to float32)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(f32)


def clamp_pad(p, hi):
    return np.minimum(np.maximum(p, f32(0.5)), f32(hi)).astype(f32)


def bp3d_coeffs(vec, vs):
    """The kernel's float32 BP3DAngle (csrc/geometry.cpp cone_bp_coeffs):
    DevConeParams with the texel-centre offset and the padding folded in."""
    nu, nv, dn = O3.cone_bp_params(vec, "default", vs)
    cu = np.array([nu[k] + 1.5 * dn[k] for k in range(4)], dtype=f32)
    cv = np.array([nv[k] + 1.5 * dn[k] for k in range(4)], dtype=f32)
    cd = np.array(dn, dtype=f32)
    return cu, cv, cd


def footprint3d(c, x0, x1, y0, y1, z0, z1, hiu, hiv, pitch, prow, margin=1):
    cu, cv, cd = c
    us, vs_, ds = [], [], []
    for j in range(4):
        x = x1 if j & 1 else x0
        y = y1 if j & 2 else y0
        un = fma(y, cu[1], fma(x, cu[0], cu[3]))
        dn = fma(y, cd[1], fma(x, cd[0], cd[3]))
        ds.append(dn)
        fr = f32(1.0) / dn
        us.append(clamp_pad(f32(un * fr), hiu))
        for z in (z0, z1):
            vn = fma(z, cv[2], fma(y, cv[1], fma(x, cv[0], cv[3])))
            vs_.append(clamp_pad(f32(vn * fr), hiv))
    if not (min(ds) > 0 or max(ds) < 0):
        return None
    u0 = max(0, int(min(us)) - margin)
    u1 = min(pitch - 2, int(max(us)) + margin)
    v0 = max(0, int(min(vs_)) - margin)
    v1 = min(prow - 1, int(max(vs_)) + margin)
    return u0, v0, u1 - u0 + 1, v1 - v0 + 1


def check3d(n, na, nu, nv, sod, odd, blocks, zs=8, margin=1, angle_step=1, plant=False):
    ang = np.linspace(0, 2 * np.pi, na, endpoint=False)
    vecs, vs = OG.normalise_3d(OG.gen_cone_projections(ang, nu, nv, sod, odd, 1.0, 1.0), "cone",
                               OG.vol_geom_3d(n, n, n))
    hiu, hiv = f32(nu + 2.5), f32(nv + 2.5)
    pitch, prow = nu + 4, nv + 4
    h = 0.5 * n - 0.5
    bad = 0
    for a in range(0, na, angle_step):
        c = bp3d_coeffs(vecs[a], vs)
        cu, cv, cd = c
        for (bx, by, bz) in blocks:
            X0, Y0, Z0 = bx * 16, by * 16, bz * zs
            fp = footprint3d(c, f32(X0 - h), f32(min(X0 + 15, n - 1) - h), f32(Y0 - h),
                             f32(min(Y0 + 15, n - 1) - h), f32(Z0 - h),
                             f32(Z0 - h) if plant else f32(Z0 + zs - 1 - h),
                             hiu, hiv, pitch, prow, margin)
            if fp is None:
                continue
            u0, v0, w, hh = fp
            X = np.arange(X0, min(X0 + 16, n))
            Y = np.arange(Y0, min(Y0 + 16, n))
            fx = (X.astype(f32) - f32(0.5 * n) + f32(0.5)).astype(f32)[None, :]
            fy = (Y.astype(f32) - f32(0.5 * n) + f32(0.5)).astype(f32)[:, None]
            fz = f32(f32(Z0) - f32(0.5 * n) + f32(0.5))
            un = fma(fz, cu[2], fma(fy, cu[1], fma(fx, cu[0], cu[3])))
            vn = fma(fz, cv[2], fma(fy, cv[1], fma(fx, cv[0], cv[3])))
            dn = fma(fz, cd[2], fma(fy, cd[1], fma(fx, cd[0], cd[3])))
            fr = (f32(1.0) / dn).astype(f32)
            pu = clamp_pad((un * fr).astype(f32), hiu)
            col = np.floor(pu).astype(int) - u0
            bad += int(np.count_nonzero((col < 0) | (col > w - 1)))
            for i in range(zs):
                pv = clamp_pad((vn * fr).astype(f32), hiv)
                row = np.floor(pv).astype(int) - v0
                bad += int(np.count_nonzero((row < 0) | (row > hh - 1)))
                vn = (vn + cv[2]).astype(f32)
    return bad


def _blocks(n, zs=8):
    nb, nz = n // 16, n // zs
    pick = sorted({0, nb // 3, nb // 2, nb - 1})
    zpick = sorted({0, nz // 4, nz // 2, nz - 1})
    return [(bx, by, bz) for bx in pick for by in pick for bz in zpick]


def test_bp3d_taps_inside_staged_footprint_512():
    """The benchmarked cone geometry (512^3, 768^2, SOD 1000, ODD 500),
    corner / edge / centre blocks at every 6th of the 720 angles."""
    assert check3d(512, 720, 768, 768, 1000.0, 500.0, _blocks(512), angle_step=6) == 0


def test_bp3d_taps_inside_staged_footprint_near_source():
    assert check3d(48, 24, 96, 96, 40.0, 400.0, _blocks(48)) == 0


def test_bp3d_planted_first_slice_footprint_is_caught():
    """A footprint from the block's first z slice only misses taps."""
    assert check3d(64, 36, 96, 96, 150.0, 100.0, _blocks(64), plant=True) > 0


def check2d(kind, w, h, nd, na, sod, odd, py, margin=1, wcap=256, plant=False):
    ang = np.linspace(0.2, 0.2 + 2 * np.pi, na, endpoint=False)
    vg = OG.vol_geom_2d(h, w)
    if kind == "par":
        vecs, _ = OG.normalise_2d(OG.gen_par_projections(ang, nd, 1.1), "parallel", vg)
    else:
        vecs, _ = OG.normalise_2d(OG.gen_fan_projections(ang, nd, sod, odd, 1.3), "fan", vg)
    pitch, hi = nd + 4, f32(nd + 2.5)
    bad = 0
    for a in range(na):
        p = [float(x) for x in vecs[a]]
        if kind == "par":
            rx, ry, dsx, dsy, ux, uy = p
            d = ux * ry - uy * rx
            num = np.array([(dsy * rx - dsx * ry) / d + 1.5, ry / d, -rx / d], dtype=f32)
            den = np.array([1, 0, 0], dtype=f32)
        else:
            sx, sy, dsx, dsy, ux, uy = p
            smdx, smdy = sx - dsx, sy - dsy
            det2 = lambda ax, ay, bx, by: ax * by - ay * bx
            sc = math.sqrt(ux * ux + uy * uy) / det2(ux, uy, smdx, smdy)
            nc, nx, ny = sc * det2(sx, sy, dsx, dsy), sc * smdy, -sc * smdx
            dc, dx, dy = sc * det2(ux, uy, sx, sy), sc * uy, -sc * ux
            num = np.array([nc + 1.5 * dc, nx + 1.5 * dx, ny + 1.5 * dy], dtype=f32)
            den = np.array([dc, dx, dy], dtype=f32)
        for X0 in range(0, w, 64):
            for Yt0 in range(0, h, 4 * py):
                cx = [f32(f32(X0) - f32(0.5 * w) + f32(0.5)),
                      f32(f32(min(X0 + 63, w - 1)) - f32(0.5 * w) + f32(0.5))]
                ylast = Yt0 if plant else min(Yt0 + 4 * py - 1, h - 1)
                cy = [-f32(f32(Yt0) - f32(0.5 * h) + f32(0.5)),
                      -f32(f32(ylast) - f32(0.5 * h) + f32(0.5))]
                ts, ds = [], []
                for j in range(4):
                    fx, fy = cx[j & 1], cy[(j >> 1) & 1]
                    nn = fma(fy, num[2], fma(fx, num[1], num[0]))
                    dd = fma(fy, den[2], fma(fx, den[1], den[0]))
                    ds.append(dd)
                    fr = f32(1.0) if kind == "par" else f32(f32(1.0) / dd)
                    ts.append(clamp_pad(f32(nn * fr), hi))
                if kind != "par" and not (min(ds) > 0 or max(ds) < 0):
                    continue
                t0 = max(0, int(min(ts)) - margin)
                t1 = min(pitch - 1, int(max(ts)) + 1 + margin)
                if t1 - t0 + 1 > wcap:
                    continue
                X = np.arange(X0, min(X0 + 64, w))
                fx = (X.astype(f32) - f32(0.5 * w) + f32(0.5)).astype(f32)
                for Y0 in range(Yt0, min(Yt0 + 4 * py, h), py):
                    fy0 = -f32(f32(Y0) - f32(0.5 * h) + f32(0.5))
                    nn = fma(fy0, num[2], fma(fx, num[1], num[0]))
                    dd = fma(fy0, den[2], fma(fx, den[1], den[0]))
                    for i in range(py):
                        if Y0 + i >= h:
                            break
                        fr = np.ones_like(dd) if kind == "par" else (f32(1.0) / dd).astype(f32)
                        t = clamp_pad((nn * fr).astype(f32), hi)
                        o = np.floor(t).astype(int) - t0
                        bad += int(np.count_nonzero((o < 0) | (o > t1 - t0 - 1)))
                        nn = (nn - num[2]).astype(f32)
                        dd = (dd - den[2]).astype(f32)
    return bad


@pytest.mark.parametrize("py", [2, 4, 8])
@pytest.mark.parametrize("case", [("par", 130, 70, 200, 45, 0, 0), ("fan", 129, 97, 190, 60, 500.0, 300.0),
                                  ("fan", 96, 96, 200, 50, 60.0, 400.0)],
                         ids=["par", "fan", "fan_near"])
def test_bp2d_taps_inside_staged_footprint(case, py):
    kind, w, h, nd, na, sod, odd = case
    assert check2d(kind, w, h, nd, na, sod, odd, py) == 0


def test_bp2d_planted_first_row_footprint_is_caught():
    assert check2d("fan", 129, 97, 190, 60, 500.0, 300.0, 4, plant=True) > 0


@pytest.mark.parametrize("case", [("par", 1024, 1024, 1536, 8, 0, 0), ("fan", 512, 512, 768, 12, 1024.0, 512.0)],
                         ids=["par1024", "fan512"])
def test_bp2d_taps_inside_staged_footprint_batched_sizes(case):
    """Slices of the batched workload's sizes (fan: source-origin 2N,
    origin-detector N), 8 rows per thread: the staged taps read LDS without
    a clamp, so every tap must stay inside its footprint."""
    kind, w, h, nd, na, sod, odd = case
    assert check2d(kind, w, h, nd, na, sod, odd, 8) == 0
