"""Memory-bounded 3D FP / BP / FDK jobs on host-resident data
(CCompositeGeometryManager::splitJobs / doJobs / doFDK,
src/CompositeGeometryManager.cpp:218-379, 1283-1304, 1409-1477).

A job runs in one piece when its GPU memory (the host data it uploads plus
the projector's internal copies) fits the budget: the limit given to
``astra.set_gpu_index(idx, memory=...)``, else the device's free memory,
times 0.9 (:1415-1431).  Otherwise the output is split, as splitJobs does
(projections over angles, the volume over z), into parts of at most a third
of the budget, and for each output part the input over its other axis so
that input part + scratch fit the rest.  Each (output part, input part)
pair runs as its own projection with the parts' geometries (a z slab of the
volume window, a subset of the projection vectors) and accumulates into the
output part, which is then written back.

The split is exact: FP of a volume is the sum of the FPs of its z slabs
(each slab samples the same world points and reads zeros outside its own
voxels, and linear interpolation is linear in the data), and the BP of a
voxel sums independent per-angle terms.  As in the reference, the input of
each part is first reduced to what can reach the output part: a BP slab
uploads only the detector rows its window (grown by half a voxel, plus a
row of margin) projects onto, and an FP skips the slices that project off
the detector (:382-410, 526-625, 854-866).  Data already on the GPU is used
in place and never split.  FDK filters the whole projection set once on
the GPU and back projects it slab by slab, so the projections (and their
padded copy) must fit.
"""
import ctypes

import numpy as np
import torch

from . import _core, _lib
from . import geometry as G
from .log import AstraError

MODE_ADD = _lib.MODE_ADD
MODE_SET = _lib.MODE_SET
_PAD = 2  # padded border of the internal copies, texels per side


def budget(dev):
    """GPU bytes one job may use on `dev`."""
    lim = _core.gpu_memory_limit()
    if not lim:
        free, _ = torch.cuda.mem_get_info(dev)
        lim = free
    return lim * 9 // 10


# --------------------------------------------------------------- sizes ----
def _fp_scratch(vshape):
    """Bytes of the FP's largest padded row-pair group copy of a volume."""
    nz, ny, nx = vshape
    groups = ((nx, ny, nz), (ny, nx, nz), (nz, nx, ny))
    return max(n0 * (n1 + 2 * _PAD) * (n2 + 2 * _PAD) for n0, n1, n2 in groups) * 8


def _bp_scratch(pshape):
    """Bytes of the BP's padded angle-major row-pair projection copy."""
    nv, na, nu = pshape
    return na * (nv + 2 * _PAD) * (nu + 2 * _PAD) * 8


def _nbytes(shape):
    return int(np.prod(shape)) * 4


def _chunks(n, per):
    per = max(1, int(per))
    return [(a, min(n, a + per)) for a in range(0, n, per)]


def plan_fp(vshape, pshape, cap, vol_on_host=True, proj_on_host=True):
    """[(angle range, [z ranges])] of a split FP, or None when it fits."""
    need = _fp_scratch(vshape) + (_nbytes(vshape) if vol_on_host else 0) + \
        (_nbytes(pshape) if proj_on_host else 0)
    if need <= cap:
        return None
    nv, na, nu = pshape
    nz, ny, nx = vshape
    per_angle = nv * nu * 4
    out_cap = cap // 3 if vol_on_host else cap // 2
    angles = _chunks(na, out_cap // per_angle)
    plan = []
    for a0, a1 in angles:
        rest = cap - (a1 - a0) * per_angle
        # slab bytes + its scratch, per slice (upper bound over the groups)
        per_slice = max(_fp_scratch((2, ny, nx)) // 2, 1) + (nx * ny * 4 if vol_on_host else 0)
        slabs = _chunks(nz, rest // per_slice)
        if rest <= 0 or (slabs and rest < per_slice):
            raise AstraError("FP job does not fit in %d bytes of GPU memory" % cap)
        plan.append(((a0, a1), slabs))
    return plan


def plan_bp(vshape, pshape, cap, vol_on_host=True, proj_on_host=True):
    """[(z range, [angle ranges])] of a split BP, or None when it fits."""
    need = _bp_scratch(pshape) + (_nbytes(vshape) if vol_on_host else 0) + \
        (_nbytes(pshape) if proj_on_host else 0)
    if need <= cap:
        return None
    nv, na, nu = pshape
    nz, ny, nx = vshape
    per_slice = nx * ny * 4
    out_cap = cap // 3 if proj_on_host else cap // 2
    slabs = _chunks(nz, out_cap // per_slice)
    per_angle = (nv + 2 * _PAD) * (nu + 2 * _PAD) * 8 + (nv * nu * 4 if proj_on_host else 0)
    plan = []
    for z0, z1 in slabs:
        rest = cap - (z1 - z0) * per_slice
        if rest < per_angle:
            raise AstraError("BP job does not fit in %d bytes of GPU memory" % cap)
        plan.append(((z0, z1), _chunks(na, rest // per_angle)))
    return plan


# ---------------------------------------------------------- geometries ----
def vol_slab_geom(vg, z0, z1):
    """Volume geometry of slices [z0, z1) (the window shrinks along z)."""
    nz = int(vg["GridSliceCount"])
    lo, hi = G._window(vg, "Z", nz)
    out = {k: v for k, v in vg.items() if k not in G.OPTION_KEYS}
    opt = G.options_of(vg)
    for ax, n in (("X", vg["GridColCount"]), ("Y", vg["GridRowCount"])):
        wlo, whi = G._window(vg, ax, n)
        opt["WindowMin" + ax], opt["WindowMax" + ax] = float(wlo), float(whi)
    dz = (float(hi) - float(lo)) / nz
    opt["WindowMinZ"] = float(lo) + z0 * dz
    opt["WindowMaxZ"] = float(lo) + z1 * dz
    out["option"] = opt
    out["GridSliceCount"] = z1 - z0
    return out


def proj_angle_geom(pg, a0, a1):
    """Projection geometry of angles [a0, a1) (vector form)."""
    v = G.geom_2vec(pg)
    v = dict(v)
    v["Vectors"] = np.asarray(v["Vectors"])[a0:a1]
    return v


def proj_rows_geom(pg, r0, r1):
    """Projection geometry (vector form) of detector rows [r0, r1): the same
    pixels, with the detector centre moved along v."""
    v = dict(G.geom_2vec(pg))
    vec = np.array(v["Vectors"], dtype=np.float64)
    nv = int(v["DetectorRowCount"])
    vec[:, 3:6] += (r0 - 0.5 * nv + 0.5 * (r1 - r0)) * vec[:, 9:12]
    v["Vectors"] = vec
    v["DetectorRowCount"] = r1 - r0
    return v


# ------------------------------------------------- vertical reduction ----
def vertical_range(vg, pg):
    """Detector rows (fractional, clipped to [-1, rows]) that the volume
    window grown by half a voxel projects onto over all projections; (0, 0)
    when it misses the detector (reduceProjectionVertical,
    src/CompositeGeometryManager.cpp:382-410)."""
    px, py, pz = (float(x) for x in G.pixel_lengths(vg))
    box = []
    for ax, n, h in (("X", vg["GridColCount"], px), ("Y", vg["GridRowCount"], py),
                     ("Z", vg["GridSliceCount"], pz)):
        lo, hi = G._window(vg, ax, n)
        box += [float(lo) - 0.5 * h, float(hi) + 0.5 * h]
    _, _, vmin, vmax = G.projected_bbox_3d(pg, *box)
    vmin = max(vmin, -1.0)
    vmax = min(vmax, float(pg["DetectorRowCount"]))
    return (0.0, 0.0) if vmin >= vmax else (vmin, vmax)


def rows_needed(vg, pg):
    """[r0, r1): the projection rows a volume part reads in a BP, one row of
    margin each side (reduceProjectionPart, :854-866); (0, 0) when none."""
    vmin, vmax = vertical_range(vg, pg)
    if vmin == vmax:  # off the detector (the reference keeps row 0 here; it adds nothing)
        return (0, 0)
    nv = int(pg["DetectorRowCount"])
    r0 = max(int(np.floor(vmin - 1.0)), 0)
    r1 = min(int(np.ceil(vmax + 1.0)), nv)
    return (0, 0) if r0 >= r1 else (r0, r1)


def slices_needed(vg, pg):
    """[z0, z1): the volume slices whose projection can reach the detector
    rows, one slice of margin each side (reduceVolumePart, :526-625, with
    its binary searches for the first and last slice); (0, 0) when the
    whole volume projects off the detector."""
    nz = int(vg["GridSliceCount"])
    full = vertical_range(vg, pg)
    if full[0] >= full[1]:
        return (0, 0)

    def off(z0, z1):  # slices [z0, z1) reach no row of `full`
        lo, hi = vertical_range(vol_slab_geom(vg, z0, z1), pg)
        return lo == hi or lo >= full[1] or hi <= full[0]

    lo, hi = 0, nz - 1
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if off(0, mid):
            lo = mid
        else:
            hi = mid - 1
    top = lo
    lo, hi = top + 1, nz
    while lo < hi:
        mid = (lo + hi) // 2
        if off(mid, nz):
            hi = mid
        else:
            lo = mid + 1
    return max(top - 1, 0), min(hi + 1, nz)


def _sub_projector(p, pg, vg):
    return _core.Projector("cuda3d", pg, vg, p.options)


# ---------------------------------------------------------------- jobs ----
def _host_array(d):
    return None if d.is_gpu() else d.numpy_view()


def fp(p, dvol, dproj, mode):
    """dproj (=|+=) FP(dvol) for a 3D projector, split when needed."""
    dev = p.device()
    vol_host, proj_host = _host_array(dvol), _host_array(dproj)
    plan = None
    if (vol_host is not None or proj_host is not None) and p.kernel != _lib.KER3D_SUM_SQUARE_WEIGHTS:
        plan = plan_fp(p.vol_shape, p.proj_shape, budget(dev), vol_host is not None,
                       proj_host is not None)
    if plan is None:
        v = dvol.on_device(dev)
        s = dproj.on_device(dev, init=(mode == MODE_ADD))
        _core.fp(p, v, s, mode)
        dproj.write_back(s)
        _core.sync(dev)
        return
    nv, na, nu = p.proj_shape
    for (a0, a1), slabs in plan:
        pg = proj_angle_geom(p.proj_geom, a0, a1)
        # only the slices that can reach this part's rows (the rest add 0)
        z0r, z1r = slices_needed(p.vol_geom, pg)
        slabs = [(z0, z1) for z0, z1 in _chunks(z1r - z0r, slabs[0][1] - slabs[0][0])]
        slabs = [(z0 + z0r, z1 + z0r) for z0, z1 in slabs]
        if proj_host is not None:
            s = (torch.from_numpy(np.ascontiguousarray(proj_host[:, a0:a1, :])).to(dev)
                 if mode == MODE_ADD else torch.zeros((nv, a1 - a0, nu), device=dev))
        else:
            full = dproj.linked
            s = full[:, a0:a1, :].contiguous() if mode == MODE_ADD else \
                torch.zeros((nv, a1 - a0, nu), device=dev)
        for z0, z1 in slabs:
            sp = _sub_projector(p, pg, vol_slab_geom(p.vol_geom, z0, z1))
            if vol_host is not None:
                v = torch.from_numpy(np.ascontiguousarray(vol_host[z0:z1])).to(dev)
            else:
                v = dvol.linked[z0:z1].contiguous()
            _core.fp(sp, v, s, MODE_ADD)
            sp.release()
        if proj_host is not None:
            proj_host[:, a0:a1, :] = s.cpu().numpy()
        else:
            dproj.linked[:, a0:a1, :].copy_(s)
        _core.sync(dev)


def bp(p, dproj, dvol, mode, kernel=None):
    """dvol (=|+=) BP(dproj) for a 3D projector, split when needed."""
    dev = p.device()
    vol_host, proj_host = _host_array(dvol), _host_array(dproj)
    plan = None
    if vol_host is not None or proj_host is not None:
        plan = plan_bp(p.vol_shape, p.proj_shape, budget(dev), vol_host is not None,
                       proj_host is not None)
    if plan is None:
        s = dproj.on_device(dev)
        v = dvol.on_device(dev, init=(mode == MODE_ADD))
        _core.bp(p, s, v, mode, kernel=kernel)
        dvol.write_back(v)
        _core.sync(dev)
        return
    nz, ny, nx = p.vol_shape
    nv = p.proj_shape[0]
    matched = (p.kernel if kernel is None else kernel) == _lib.KER3D_MATCHED_BP
    for (z0, z1), chunks in plan:
        vg = vol_slab_geom(p.vol_geom, z0, z1)
        if vol_host is not None:
            v = (torch.from_numpy(np.ascontiguousarray(vol_host[z0:z1])).to(dev)
                 if mode == MODE_ADD else torch.zeros((z1 - z0, ny, nx), device=dev))
        else:
            v = dvol.linked[z0:z1].clone() if mode == MODE_ADD else \
                torch.zeros((z1 - z0, ny, nx), device=dev)
        for a0, a1 in chunks:
            pg = proj_angle_geom(p.proj_geom, a0, a1)
            if matched:  # the transpose of the FP keeps the FP's full detector
                r0, r1 = 0, nv
                sp = _sub_projector(p, pg, vg)
            else:
                r0, r1 = rows_needed(vg, pg)  # rows outside add nothing to this slab
                if r0 == r1:
                    continue
                sp = _sub_projector(p, proj_rows_geom(pg, r0, r1), vg)
            if proj_host is not None:
                s = torch.from_numpy(np.ascontiguousarray(proj_host[r0:r1, a0:a1, :])).to(dev)
            else:
                s = dproj.linked[r0:r1, a0:a1, :].contiguous()
            _core.bp(sp, s, v, MODE_ADD, kernel=kernel)
            sp.release()
        if vol_host is not None:
            vol_host[z0:z1] = v.cpu().numpy()
        else:
            dvol.linked[z0:z1].copy_(v)
        _core.sync(dev)


def fdk(p, dproj, dvol, mode, filter_handle, short_scan):
    """dvol (=|+=) FDK(dproj); dproj is left unchanged.  The volume is split
    over z when it does not fit beside the filtered projections.
    `filter_handle` is a native filter owned by the caller (None: the
    library's Ram-Lak)."""
    dev = p.device()
    vol_host = _host_array(dvol)
    pt = dproj.on_device(dev).clone()
    nv, na, nu = p.proj_shape
    cap = budget(dev)
    fixed = _nbytes(p.proj_shape) + _bp_scratch(p.proj_shape)
    nz, ny, nx = p.vol_shape
    slabs = [(0, nz)]
    if vol_host is not None and fixed + _nbytes(p.vol_shape) > cap:
        if cap - fixed < nx * ny * 4:
            raise AstraError("FDK job does not fit in %d bytes of GPU memory" % cap)
        slabs = _chunks(nz, (cap - fixed) // (nx * ny * 4))
    with torch.cuda.device(dev):
        h = ctypes.c_void_p(p.handle(dev))
        _lib.call("astra_amd_fdk_filter", h, _core.ptr(pt), 1 if short_scan else 0, filter_handle,
                  _core.stream_ptr(dev))
    for z0, z1 in slabs:
        whole = (z0, z1) == (0, nz)
        sp = p if whole else _sub_projector(p, p.proj_geom, vol_slab_geom(p.vol_geom, z0, z1))
        if vol_host is not None:
            v = (torch.from_numpy(np.ascontiguousarray(vol_host[z0:z1])).to(dev)
                 if mode == MODE_ADD else torch.empty((z1 - z0, ny, nx), device=dev))
        else:
            v = dvol.on_device(dev, init=(mode == MODE_ADD))
        _core.bp(sp, pt, v, mode, kernel=_lib.KER3D_FDK_WEIGHTING)
        if vol_host is not None:
            vol_host[z0:z1] = v.cpu().numpy()
        else:
            dvol.write_back(v)
        if not whole:
            sp.release()
        _core.sync(dev)
