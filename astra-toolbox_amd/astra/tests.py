"""Basic self-checks (python/astra/tests.py): the forward projection of a
two-row image / two-slice volume at angle 0 equals its column sums.  This
build has no CPU projectors, so only the GPU checks run."""
import numpy as np


def _par2d_fp_ok():
    import astra
    vg = astra.create_vol_geom(2, 32)
    pg = astra.create_proj_geom("parallel", 1, 32, [0])
    pid = astra.create_projector("cuda", pg, vg)
    vol = np.random.rand(2, 32)
    sid, sino = astra.create_sino(vol, pid)
    astra.data2d.delete(sid)
    astra.projector.delete(pid)
    return np.max(np.abs(sino[0, :] - vol.sum(axis=0))) < 1e-6


def _par3d_fp_ok():
    import astra
    vg = astra.create_vol_geom(2, 32, 32)
    pg = astra.create_proj_geom("parallel3d", 1, 1, 32, 32, [0])
    vol = np.random.rand(32, 2, 32)
    sid, sino = astra.create_sino3d_gpu(vol, pg, vg)
    astra.data3d.delete(sid)
    return np.max(np.abs(sino[:, 0, :] - vol.sum(axis=1))) < 1e-6


def _check(label, fn):
    print("Testing basic %s functionality... " % label, end="")
    ok = fn()
    print("Ok" if ok else "Error")
    return ok


def test_noCUDA():
    """The checks that need no GPU: none in this build (it has no CPU
    projectors)."""
    import astra
    print("ASTRA Toolbox v%s" % (astra.__version__,))
    print("No CPU projectors in this build: nothing to test without a GPU")


def test_CUDA():
    """Basic GPU checks of the 2D and 3D projectors."""
    import astra
    print("ASTRA Toolbox v%s" % (astra.__version__,))
    print("Getting GPU info... ", end="")
    print(astra.get_gpu_info())
    ok2 = _check("GPU 2D", _par2d_fp_ok)
    ok3 = _check("GPU 3D", _par3d_fp_ok)
    if not (ok2 and ok3):
        raise RuntimeError("Test failed")


def test():
    """Run the basic checks this machine supports."""
    import astra
    if astra.use_cuda():
        test_CUDA()
    else:
        print("No GPU support available")
        test_noCUDA()
