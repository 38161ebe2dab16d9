"""Direct FP/BP on DLPack tensors and composite jobs
(python/astra/experimental.pyx:94-341).

Inputs and outputs may be numpy arrays, torch CPU/GPU tensors or any DLPack
producer; GPU tensors are used in place (zero copy), host arrays are
staged through the projector's GPU.  Calls are synchronous, like the
reference's CompositeGeometryManager jobs.
"""
from . import _core, _lib, composite
from . import geometry as G
from ._registry import data3d as _d3
from ._registry import projectors2d as _p2, projectors3d as _p3
from .log import AstraError

MODE_ADD = _lib.MODE_ADD
MODE_SET = _lib.MODE_SET


def _run(p, dvol, dproj, mode, t):
    if p.is3d:  # CCompositeGeometryManager jobs: split when they do not fit
        if t == "FP":
            composite.fp(p, dvol, dproj, mode)
        elif t == "BP":
            composite.bp(p, dproj, dvol, mode)
        else:
            raise AstraError("Internal error: wrong op type")
        return
    dev = p.device()
    if t == "FP":
        v = dvol.on_device(dev)
        s = dproj.on_device(dev, init=(mode == MODE_ADD))
        _core.fp(p, v, s, mode)
        dproj.write_back(s)
    elif t == "BP":
        s = dproj.on_device(dev)
        v = dvol.on_device(dev, init=(mode == MODE_ADD))
        _core.bp(p, s, v, mode)
        dvol.write_back(v)
    else:
        raise AstraError("Internal error: wrong op type")
    _core.sync(dev)


def direct_FPBP3D(projector_id, vol, proj, mode, t):
    if mode not in (MODE_ADD, MODE_SET):
        raise AstraError("Internal error: wrong composite mode")
    p = _p3.get(projector_id)
    dv = _core.link_data("-vol", p.vol_geom, p.vol_shape, vol)
    dp = _core.link_data("-sino", p.proj_geom, p.proj_shape, proj)
    _run(p, dv, dp, mode, t)


def direct_FP3D(projector_id, vol, proj):
    """3D forward projection into the pre-allocated `proj` (MODE_SET)."""
    direct_FPBP3D(projector_id, vol, proj, MODE_SET, "FP")


def direct_BP3D(projector_id, vol, proj):
    """3D back projection of `proj` into the pre-allocated `vol` (MODE_SET)."""
    direct_FPBP3D(projector_id, vol, proj, MODE_SET, "BP")


def direct_FPBP2D(projector_id, vol, proj, t):
    p = _p2.get(projector_id)
    dv = _core.link_data("-vol", p.vol_geom, p.vol_shape, vol)
    dp = _core.link_data("-sino", p.proj_geom, p.proj_shape, proj)
    _run(p, dv, dp, MODE_SET, t)


def direct_FP2D(projector_id, vol, proj):
    direct_FPBP2D(projector_id, vol, proj, "FP")


def direct_BP2D(projector_id, vol, proj):
    direct_FPBP2D(projector_id, vol, proj, "BP")


def do_composite(projector_id, vol_ids, proj_ids, mode, t):
    """FP / BP between lists of 3D data objects
    (CCompositeGeometryManager::doFP/doBP): every projection receives the
    contribution of every volume (FP), every volume that of every projection
    (BP).  One GPU; each job is split when it does not fit."""
    if mode not in (MODE_ADD, MODE_SET):
        raise AstraError("Internal error: wrong composite mode")
    base = _p3.get(projector_id) if projector_id is not None and projector_id in _p3.objects else None
    opts = dict(base.options) if base is not None else {}
    vols = [_d3.get(i) for i in vol_ids]
    projs = [_d3.get(i) for i in proj_ids]
    outs, ins = (projs, vols) if t == "FP" else (vols, projs)
    for o in outs:
        first = True
        for i in ins:
            pv, pp = (i, o) if t == "FP" else (o, i)
            p = _core.projector_for(pp.geom, pv.geom, opts)
            m = mode if first else MODE_ADD
            _run(p, pv, pp, m, t)
            first = False


def do_composite_FP(projector_id, vol_ids, proj_ids):
    do_composite(projector_id, vol_ids, proj_ids, MODE_SET, "FP")


def do_composite_BP(projector_id, vol_ids, proj_ids):
    do_composite(projector_id, vol_ids, proj_ids, MODE_SET, "BP")


def accumulate_FP(projector_id, vol_id, proj_id):
    """proj += FP(vol) between two 3D data objects (MODE_ADD composite job)."""
    do_composite(projector_id, [vol_id], [proj_id], MODE_ADD, "FP")


def accumulate_BP(projector_id, vol_id, proj_id):
    """vol += BP(proj) between two 3D data objects (MODE_ADD composite job)."""
    do_composite(projector_id, [vol_id], [proj_id], MODE_ADD, "BP")


def getProjectedBBox(geometry, minx, maxx, miny, maxy, minz, maxz):
    """(vmin, vmax): the detector rows that the box [minx, maxx] x [miny,
    maxy] x [minz, maxz] projects onto, over all projections of a 3D
    projection geometry (row index coordinates, 0 = centre of the first
    row)."""
    _, _, vmin, vmax = G.projected_bbox_3d(geometry, minx, maxx, miny, maxy, minz, maxz)
    return (vmin, vmax)


def projectPoint(geometry, x, y, z, angle):
    """(u, v) detector index coordinates of world point (x, y, z) in
    projection number `angle` of a 3D projection geometry."""
    n = G.proj_shape(geometry)[1]
    if not 0 <= int(angle) < n:
        raise AstraError("projection index out of range")
    return G.project_point_3d(geometry, x, y, z, int(angle))


def accumulate_FDK(projector_id, vol_id, proj_id):
    """vol += FDK(proj) with a Ram-Lak filter (CCompositeGeometryManager::doFDK,
    MODE_ADD); the projection data object is left unchanged."""
    p = _p3.get(projector_id)
    dv, dp = _d3.get(vol_id), _d3.get(proj_id)
    composite.fdk(p, dp, dv, MODE_ADD, None, False)
