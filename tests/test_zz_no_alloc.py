"""astra_amd_fp2d_noalloc: the 2D forward projection whose volume copies
live in a fixed per-GPU array.  Called through the C ABI and compared bit
for bit with astra_amd_fp2d (same kernels, same copies), on two calls in a
row, the second on a smaller volume, and checked to take nothing from the
stream-ordered memory pool."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _geom(kind, w, h, na, nd):
    import astra
    vg = astra.create_vol_geom(h, w)
    ang = np.linspace(0.1, 0.1 + 2 * np.pi, na, endpoint=False)
    if kind == "fan":
        pg = astra.create_proj_geom("fanflat", 1.25, nd, ang, 700.0, 300.0)
    else:
        pg = astra.create_proj_geom("parallel", 0.9, nd, ang)
    from astra import _core
    return _core.Projector("cuda", pg, vg, {"DetectorSuperSampling": 2 if kind == "fan" else 1})


def _run(entry, proj, vol_t, sino_t, mode):
    import torch
    from astra import _core, _lib
    dev = vol_t.device
    h = ctypes.c_void_p(proj.handle(dev))
    p = proj.params2d(1.0)
    with torch.cuda.device(dev):
        _lib.call(entry, h, _core.ptr(vol_t), _core.ptr(sino_t), ctypes.byref(p), mode,
                  _core.stream_ptr(dev))
        torch.cuda.synchronize(dev)


def _pool():
    from astra import _lib
    cur, peak = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.call("astra_amd_scratch_usage", ctypes.byref(cur), ctypes.byref(peak), 1)
    return cur.value, peak.value


@pytest.mark.parametrize("case", [0])
def test_noalloc_equals_fp2d_on_two_calls(case):
    import torch
    from astra import _lib
    rng = np.random.default_rng(11)
    calls = [("fan", 700, 530, 240, 900, _lib.MODE_SET),
             ("parallel", 190, 250, 120, 300, _lib.MODE_ADD)]
    for kind, w, h, na, nd, mode in calls:
        print("noalloc call: %s %dx%d, %d angles x %d dets, mode %d" % (kind, w, h, na, nd, mode),
              flush=True)
        proj = _geom(kind, w, h, na, nd)
        vol = torch.from_numpy(rng.random((h, w), dtype=np.float32)).cuda()
        start = torch.from_numpy(rng.random((na, nd), dtype=np.float32)).cuda()
        ref = start.clone()
        got = start.clone()
        _run("astra_amd_fp2d", proj, vol, ref, mode)
        torch.cuda.synchronize()
        cur0, _ = _pool()  # also resets the pool's high-water mark
        _run("astra_amd_fp2d_noalloc", proj, vol, got, mode)
        cur1, peak1 = _pool()
        assert cur1 == cur0 and peak1 == cur0, (cur0, cur1, peak1)
        assert float(ref.abs().max()) > 0
        assert torch.equal(got, ref), float((got - ref).abs().max())
        proj.release()


def test_noalloc_refuses_a_volume_beyond_the_array():
    import torch
    from astra import _lib
    from astra.log import AstraError
    proj = _geom("parallel", 3000, 3000, 4, 8)  # 2 x 3000 x 3004 floats > 16 Mi
    small = torch.zeros(16, dtype=torch.float32, device="cuda")
    with pytest.raises(AstraError, match="fp2d_noalloc"):
        _run("astra_amd_fp2d_noalloc", proj, small, small, _lib.MODE_SET)
    proj.release()
