"""Geometry / object creation helpers (python/astra/creators.py:36-607)."""
import numpy as np

from . import algorithm, data2d, data3d, projector, projector3d
from .log import AstraError


def astra_dict(intype):
    """Configuration dictionary skeleton for an object type."""
    if intype == "SIRT_CUDA2":
        intype = "SIRT_CUDA"
    elif intype == "FP_CUDA2":
        intype = "FP_CUDA"
    return {"type": intype}


def create_vol_geom(*args):
    """2D: (N) | ((rows, cols)) | (rows, cols) | (rows, cols, minx, maxx, miny, maxy);
    3D: ((rows, cols, slices)) | (rows, cols, slices) |
    (rows, cols, slices, minx, maxx, miny, maxy, minz, maxz)."""
    if len(args) == 1 and isinstance(args[0], (int, np.integer)):
        dims, win = (args[0], args[0]), ()
    elif len(args) == 1:
        dims, win = tuple(args[0]), ()
    elif len(args) in (2, 3):
        dims, win = tuple(args), ()
    elif len(args) == 6:
        dims, win = tuple(args[:2]), tuple(args[2:])
    elif len(args) == 9:
        dims, win = tuple(args[:3]), tuple(args[3:])
    else:
        raise AstraError("create_vol_geom: unsupported arguments")
    if len(dims) not in (2, 3):
        raise AstraError("create_vol_geom: expected 2 or 3 dimensions")
    geom = {"GridRowCount": dims[0], "GridColCount": dims[1], "option": {}}
    if len(dims) == 3:
        geom["GridSliceCount"] = dims[2]
    axes = ["X", "Y", "Z"][: len(dims)]
    counts = {"X": dims[1], "Y": dims[0], "Z": dims[2] if len(dims) == 3 else None}
    for k, ax in enumerate(axes):
        if win:
            lo, hi = win[2 * k], win[2 * k + 1]
        else:
            lo, hi = -counts[ax] / 2.0, counts[ax] / 2.0
        geom["option"]["WindowMin" + ax] = lo
        geom["option"]["WindowMax" + ax] = hi
    return geom


_PROJ_ARGS = {
    "parallel": ("DetectorWidth", "DetectorCount", "ProjectionAngles"),
    "parallel_vec": ("DetectorCount", "Vectors"),
    "fanflat": ("DetectorWidth", "DetectorCount", "ProjectionAngles", "DistanceOriginSource",
                "DistanceOriginDetector"),
    "fanflat_vec": ("DetectorCount", "Vectors"),
    "parallel3d": ("DetectorSpacingX", "DetectorSpacingY", "DetectorRowCount", "DetectorColCount",
                   "ProjectionAngles"),
    "cone": ("DetectorSpacingX", "DetectorSpacingY", "DetectorRowCount", "DetectorColCount",
             "ProjectionAngles", "DistanceOriginSource", "DistanceOriginDetector"),
    "cone_vec": ("DetectorRowCount", "DetectorColCount", "Vectors"),
    "parallel3d_vec": ("DetectorRowCount", "DetectorColCount", "Vectors"),
    "cyl_cone_vec": ("DetectorRowCount", "DetectorColCount", "Vectors"),
    "sparse_matrix": ("DetectorWidth", "DetectorCount", "ProjectionAngles", "MatrixID"),
}
_VEC_COLS = {"parallel_vec": 6, "fanflat_vec": 6, "cone_vec": 12, "parallel3d_vec": 12,
             "cyl_cone_vec": 13}


def create_proj_geom(intype, *args):
    """Projection geometry dictionary, same argument order as the reference."""
    if intype not in _PROJ_ARGS:
        raise AstraError("Unknown projection geometry type: " + str(intype))
    keys = _PROJ_ARGS[intype]
    if len(args) < len(keys):
        raise AstraError("Not enough variables. Usage: astra.create_proj_geom(%s, %s)"
                         % (intype, ", ".join(keys)))
    geom = {"type": intype}
    geom.update(zip(keys, args))
    if intype in _VEC_COLS:
        v = np.asarray(geom["Vectors"])
        if v.ndim != 2 or v.shape[1] != _VEC_COLS[intype]:
            raise AstraError("V should be a Nx%d matrix, with N the number of projections"
                             % _VEC_COLS[intype])
    return geom


def create_projector(proj_type, proj_geom, vol_geom, options=None):
    cfg = astra_dict(proj_type)
    cfg["ProjectionGeometry"] = proj_geom
    cfg["VolumeGeometry"] = vol_geom
    if options is not None:
        cfg["options"] = options
    if proj_type == "cuda3d":
        return projector3d.create(cfg)
    return projector.create(cfg)


def _run_once(cfg, iterations=1):
    aid = algorithm.create(cfg)
    try:
        algorithm.run(aid, iterations)
    finally:
        algorithm.delete(aid)


def create_sino(data, proj_id, returnData=True, gpuIndex=None):
    pg = projector.projection_geometry(proj_id)
    vg = projector.volume_geometry(proj_id)
    own = isinstance(data, np.ndarray)
    vid = data2d.create("-vol", vg, data) if own else data
    sid = data2d.create("-sino", pg, 0)
    cfg = astra_dict("FP_CUDA")
    cfg.update(ProjectorId=proj_id, ProjectionDataId=sid, VolumeDataId=vid)
    if gpuIndex is not None:
        cfg["option"] = {"GPUindex": gpuIndex}
    _run_once(cfg)
    if own:
        data2d.delete(vid)
    return (sid, data2d.get(sid)) if returnData else sid


def create_backprojection(data, proj_id, returnData=True):
    pg = projector.projection_geometry(proj_id)
    vg = projector.volume_geometry(proj_id)
    own = isinstance(data, np.ndarray)
    sid = data2d.create("-sino", pg, data) if own else data
    vid = data2d.create("-vol", vg, 0)
    cfg = astra_dict("BP_CUDA")
    cfg.update(ProjectorId=proj_id, ProjectionDataId=sid, ReconstructionDataId=vid)
    _run_once(cfg)
    if own:
        data2d.delete(sid)
    return (vid, data2d.get(vid)) if returnData else vid


def create_sino3d_gpu(data, proj_geom, vol_geom, returnData=True, gpuIndex=None):
    own = isinstance(data, np.ndarray)
    vid = data3d.create("-vol", vol_geom, data) if own else data
    sid = data3d.create("-sino", proj_geom, 0)
    cfg = astra_dict("FP3D_CUDA")
    cfg.update(ProjectionDataId=sid, VolumeDataId=vid)
    if gpuIndex is not None:
        cfg["option"] = {"GPUindex": gpuIndex}
    _run_once(cfg)
    if own:
        data3d.delete(vid)
    return (sid, data3d.get(sid)) if returnData else sid


def create_backprojection3d_gpu(data, proj_geom, vol_geom, returnData=True):
    own = isinstance(data, np.ndarray)
    sid = data3d.create("-sino", proj_geom, data) if own else data
    vid = data3d.create("-vol", vol_geom, 0)
    cfg = astra_dict("BP3D_CUDA")
    cfg.update(ProjectionDataId=sid, ReconstructionDataId=vid)
    _run_once(cfg)
    if own:
        data3d.delete(sid)
    return (vid, data3d.get(vid)) if returnData else vid


def create_reconstruction(rec_type, proj_id, sinogram, iterations=1, use_mask="no",
                          mask=np.array([]), use_minc="no", minc=0, use_maxc="no", maxc=255,
                          returnData=True, filterType=None, filterData=None):
    pg = projector.projection_geometry(proj_id)
    vg = projector.volume_geometry(proj_id)
    own = isinstance(sinogram, np.ndarray)
    sid = data2d.create("-sino", pg, sinogram) if own else sinogram
    rid = data2d.create("-vol", vg, 0.0)
    cfg = astra_dict(rec_type)
    cfg.update(ProjectorId=proj_id, ProjectionDataId=sid, ReconstructionDataId=rid)
    cfg["options"] = {}
    mid = None
    if use_mask == "yes":
        mid = data2d.create("-vol", vg, mask) if isinstance(mask, np.ndarray) else mask
        cfg["options"]["ReconstructionMaskId"] = mid
    if filterType is not None:
        cfg["FilterType"] = filterType
    if filterData is not None:
        raise AstraError("custom filter data is not part of this build")
    cfg["options"]["UseMinConstraint"] = use_minc
    cfg["options"]["MinConstraintValue"] = minc
    cfg["options"]["UseMaxConstraint"] = use_maxc
    cfg["options"]["MaxConstraintValue"] = maxc
    _run_once(cfg, iterations)
    if own:
        data2d.delete(sid)
    if use_mask == "yes" and isinstance(mask, np.ndarray):
        data2d.delete(mid)
    return (rid, data2d.get(rid)) if returnData else rid
