"""The example plugins (SIRT-PLUGIN, CGLS-PLUGIN: numpy on OpTomo, every
product a GPU FP or BP) against the native algorithms on the same data.
Both compute the same iteration in float32 with different orders of
summation (host numpy vs fused GPU epilogues, FP64 vs FP32 dot products),
so they agree to rounding: 4.7e-8 (SIRT 2D), 4.9e-8 (SIRT 3D) and
1.7e-7 (CGLS 2D) rel-L2 after 10 iterations, held to 1e-5."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _problem2d():
    import astra
    vg = astra.create_vol_geom(48, 40)
    pg = astra.create_proj_geom("fanflat", 1.0, 72, np.linspace(0, 2 * np.pi, 60, False),
                                300.0, 200.0)
    pid = astra.create_projector("cuda", pg, vg)
    x = np.random.default_rng(5).random((48, 40), dtype=np.float32)
    sid, _ = astra.create_sino(x, pid)
    return astra, astra.data2d, vg, pid, sid


def _problem3d():
    import astra
    vg = astra.create_vol_geom(24, 20, 16)
    pg = astra.create_proj_geom("cone", 1.0, 1.0, 20, 32, np.linspace(0, 2 * np.pi, 36, False),
                                200.0, 100.0)
    pid = astra.create_projector("cuda3d", pg, vg)
    x = np.random.default_rng(6).random((16, 24, 20), dtype=np.float32)
    sid, _ = astra.create_sino3d_gpu(x, pg, vg)
    return astra, astra.data3d, vg, pid, sid


def _run(astra, mod, vg, pid, sid, typ, opts, its):
    rid = mod.create("-vol", vg, 0.0)
    cfg = astra.astra_dict(typ)
    cfg.update(ProjectorId=pid, ProjectionDataId=sid, ReconstructionDataId=rid, option=dict(opts))
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid, its)
    out = mod.get(rid)
    astra.algorithm.delete(aid)
    mod.delete(rid)
    return out


@pytest.mark.parametrize("case", ["sirt2d", "sirt3d", "cgls2d"])
def test_example_plugins_match_native(case):
    import astra
    astra.plugin.register(astra.plugins.SIRTPlugin)
    astra.plugin.register(astra.plugins.CGLSPlugin)
    a, mod, vg, pid, sid = _problem3d() if case == "sirt3d" else _problem2d()
    if case.startswith("sirt"):
        native = "SIRT3D_CUDA" if case == "sirt3d" else "SIRT_CUDA"
        opts = {"Relaxation": 0.8, "MinConstraint": 0.0}
        got = _run(a, mod, vg, pid, sid, "SIRT-PLUGIN", opts, 10)
        want = _run(a, mod, vg, pid, sid, native, opts, 10)
    else:
        got = _run(a, mod, vg, pid, sid, "CGLS-PLUGIN", {}, 10)
        want = _run(a, mod, vg, pid, sid, "CGLS_CUDA", {}, 10)
    e = rel_l2(got, want)
    print("%s: plugin vs native rel-L2 %.3g" % (case, e))
    assert np.abs(want).max() > 0 and e <= 1e-5


def test_reconstruction_algorithm_bases():
    """plugin.ReconstructionAlgorithm2D / 3D set pid, s, v, vg, pg before
    initialize; 2D data that does not match the projector is refused."""
    import astra

    class Probe2D(astra.plugin.ReconstructionAlgorithm2D):
        astra_name = "PROBE-2D-TEST"

        def initialize(self, cfg, Step=0.5):
            self.step = Step

        def run(self, its):
            W = astra.OpTomo(self.pid)
            for _ in range(its):
                self.v += np.float32(self.step) * W.BP(self.s - W.FP(self.v)) / np.float32(1e4)

    class Probe3D(astra.plugin.ReconstructionAlgorithm3D):
        astra_name = "PROBE-3D-TEST"

        def initialize(self, cfg):
            pass

        def run(self, its):
            self.v[...] = 1.0

    astra.plugin.register(Probe2D)
    astra.plugin.register(Probe3D)
    a, mod, vg, pid, sid = _problem2d()
    rid = mod.create("-vol", vg, 0.0)
    aid = a.algorithm.create({"type": "PROBE-2D-TEST", "ProjectorId": pid, "ProjectionDataId": sid,
                              "ReconstructionDataId": rid})
    obj = a.algorithm.get_plugin_object(aid)
    assert obj.s.shape == (60, 72) and obj.v.shape == (48, 40) and obj.step == 0.5
    assert obj.vg["GridRowCount"] == 48 and obj.pg["type"] == "fanflat"
    a.algorithm.run(aid, 2)
    assert np.abs(mod.get(rid)).max() > 0
    bad = mod.create("-vol", a.create_vol_geom(10, 10), 0.0)
    with pytest.raises(a.AstraError, match="not compatible"):
        a.algorithm.create({"type": "PROBE-2D-TEST", "ProjectorId": pid, "ProjectionDataId": sid,
                            "ReconstructionDataId": bad})
    a.algorithm.delete(aid)
    mod.delete([rid, bad])
    a3, mod3, vg3, pid3, sid3 = _problem3d()
    rid3 = mod3.create("-vol", vg3, 0.0)
    aid3 = a3.algorithm.create({"type": "PROBE-3D-TEST", "ProjectorId": pid3,
                                "ProjectionDataId": sid3, "ReconstructionDataId": rid3})
    o3 = a3.algorithm.get_plugin_object(aid3)
    assert o3.s.shape == (20, 36, 32) and o3.vg["GridSliceCount"] == 16
    a3.algorithm.run(aid3, 1)
    assert np.array_equal(mod3.get(rid3), np.ones((16, 24, 20), np.float32))
    a3.algorithm.delete(aid3)
    mod3.delete(rid3)
