"""astra.OpTomo (python/astra/optomo.py; tests/python/unit/test_optomo.py)
on the GPU projectors: products equal the data-object paths bit for bit,
flattened and shaped inputs agree, ``out`` is filled in place, the
transpose is the back projection, ``reconstruct`` equals the algorithm
API, and scipy's lsqr runs on the operator and lowers the residual."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(dim):
    import astra
    from astra import phantom
    if dim == 2:
        vg = astra.create_vol_geom(48, 56)
        pg = astra.create_proj_geom("fanflat", 1.0, 72, np.linspace(0, 2 * np.pi, 60, endpoint=False),
                                    150.0, 80.0)
        pid = astra.create_projector("cuda", pg, vg)
        vol = phantom.ellipses_2d(48, 56, seed=31)
    else:
        vg = astra.create_vol_geom(24, 28, 20)
        pg = astra.create_proj_geom("cone", 1.0, 1.0, 30, 40, np.linspace(0, 2 * np.pi, 36, endpoint=False),
                                    120.0, 60.0)
        pid = astra.create_projector("cuda3d", pg, vg)
        vol = phantom.ellipsoids_3d(28, seed=32)[:20, 2:26, :]
        vol = np.ascontiguousarray(vol)
    return astra, vg, pg, pid, vol


@pytest.mark.parametrize("dim", [2, 3])
def test_products_equal_data_object_paths(dim):
    astra, vg, pg, pid, vol = _case(dim)
    W = astra.OpTomo(pid)
    if dim == 2:
        sid, ref = astra.create_sino(vol, pid)
        bid, bref = astra.create_backprojection(ref, pid)
        mod = astra.data2d
    else:
        sid, ref = astra.create_sino3d_gpu(vol, pg, vg)
        bid, bref = astra.create_backprojection3d_gpu(ref, pg, vg)
        mod = astra.data3d
    mod.delete([sid, bid])
    assert W.shape == (ref.size, vol.size) and W.T.shape == (vol.size, ref.size)
    assert W.T.T is W
    assert np.array_equal(W.FP(vol), ref)
    assert np.array_equal(W.FP(vol.ravel().astype(np.float64)), ref)
    out = np.full(ref.shape, 5.0, np.float32)
    assert W.FP(vol, out=out) is out and np.array_equal(out, ref)
    assert np.array_equal(W * vol, ref.ravel())
    assert np.array_equal(W.matvec(vol.ravel()), ref.ravel())
    assert np.array_equal(W.BP(ref), bref)
    assert np.array_equal(W.T * ref, bref.ravel())
    assert np.array_equal(W.rmatvec(ref.ravel()), bref.ravel())
    assert np.array_equal(W.T.matvec(ref.ravel()), bref.ravel())
    outb = np.zeros(bref.shape, np.float32)
    W.BP(ref.ravel(), out=outb)
    assert np.array_equal(outb, bref)
    astra.projector.delete(pid) if dim == 2 else astra.projector3d.delete(pid)


@pytest.mark.parametrize("dim", [2, 3])
def test_reconstruct_equals_algorithm_api(dim):
    astra, vg, pg, pid, vol = _case(dim)
    W = astra.OpTomo(pid)
    sino = W.FP(vol)
    method = "SIRT_CUDA" if dim == 2 else "SIRT3D_CUDA"
    got = W.reconstruct(method, sino, 12, extraOptions={"MinConstraint": 0.0})
    mod = astra.data2d if dim == 2 else astra.data3d
    sid = mod.create("-sino", pg, sino)
    rid = mod.create("-vol", vg, 0.0)
    cfg = astra.astra_dict(method)
    cfg.update(ProjectionDataId=sid, ReconstructionDataId=rid, ProjectorId=pid,
               option={"MinConstraint": 0.0})
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid, 12)
    assert np.array_equal(got, mod.get(rid))
    astra.algorithm.delete(aid)
    mod.delete([sid, rid])


@pytest.mark.parametrize("dim", [2, 3])
def test_scipy_lsqr_on_the_operator(dim):
    import scipy.sparse.linalg as sla
    astra, vg, pg, pid, vol = _case(dim)
    W = astra.OpTomo(pid)
    sino = W.FP(vol)
    x = sla.lsqr(W, sino.ravel(), iter_lim=15)[0]
    assert x.shape == (vol.size,) and np.all(np.isfinite(x))
    r0 = np.linalg.norm(sino)
    r = np.linalg.norm(W.FP(x.astype(np.float32)) - sino)
    assert r < 0.2 * r0
