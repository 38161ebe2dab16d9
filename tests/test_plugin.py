"""Python algorithm plugins (python/astra/plugin.py and the registry of
python/astra/src/PythonPluginAlgorithmFactory.cpp:108-230), host side: a
plugin that touches only numpy data, registered by class, by
"module.Class" and under another name; options mapped onto initialize's
signature; docstrings as help."""
import numpy as np
import pytest

import astra
from astra import plugin


class Scaler(plugin.base):
    """Multiplies the reconstruction by Factor and adds Offset, per iteration."""

    astra_name = "SCALER-TEST"

    def initialize(self, cfg, Factor, Offset=0.0):
        self.vid = cfg["ReconstructionDataId"]
        self.factor, self.offset = Factor, Offset

    def run(self, its):
        v = astra.data2d.get_shared(self.vid)
        for _ in range(its):
            v *= self.factor
            v += self.offset


def _vol(value):
    return astra.data2d.create("-vol", astra.create_vol_geom(4, 5), value)


def test_register_create_run():
    plugin.register(Scaler)
    assert plugin.get_registered()["SCALER-TEST"] is Scaler
    vid = _vol(1.0)
    cfg = {"type": "SCALER-TEST", "ReconstructionDataId": vid, "options": {"Factor": 2.0}}
    aid = astra.algorithm.create(cfg)
    astra.algorithm.run(aid, 3)
    assert np.array_equal(astra.data2d.get(vid), np.full((4, 5), 8.0, np.float32))
    obj = astra.algorithm.get_plugin_object(aid)
    assert isinstance(obj, Scaler) and obj.factor == 2.0 and obj.offset == 0.0
    with pytest.raises(astra.AstraError):
        astra.algorithm.get_res_norm(aid)
    astra.algorithm.delete(aid)
    astra.data2d.delete(vid)


def test_options_and_errors():
    plugin.register(Scaler)
    vid = _vol(1.0)
    with pytest.raises(astra.AstraError, match="Missing required options"):
        astra.algorithm.create({"type": "SCALER-TEST", "ReconstructionDataId": vid})
    aid = astra.algorithm.create({"type": "SCALER-TEST", "ReconstructionDataId": vid,
                                  "options": {"Factor": 1.0, "Offset": 0.5, "Unused": 1}})
    astra.algorithm.run(aid, 2)
    assert np.array_equal(astra.data2d.get(vid), np.full((4, 5), 2.0, np.float32))
    with pytest.raises(astra.AstraError):
        astra.algorithm.create({"type": "NO-SUCH-PLUGIN", "ReconstructionDataId": vid})
    sid = astra.data2d.create("-vol", astra.create_vol_geom(2, 2), 0.0)
    with pytest.raises(astra.AstraError):
        astra.algorithm.get_plugin_object(sid)  # an id that is no plugin algorithm
    astra.algorithm.delete(aid)
    astra.data2d.delete([vid, sid])


def test_register_by_string_and_name_and_help():
    plugin.register("test_plugin.Scaler")
    assert plugin.get_registered()["SCALER-TEST"] is Scaler
    plugin.register("test_plugin.Scaler", "LAZY-SCALER")  # imported at create
    assert plugin.get_registered()["LAZY-SCALER"] == "test_plugin.Scaler"
    plugin.register(Scaler, "OTHER-NAME")
    assert plugin.get_help("LAZY-SCALER").startswith("Multiplies the reconstruction")
    assert plugin.get_help("OTHER-NAME") == plugin.get_help("SCALER-TEST")
    assert plugin.get_help("NOT-REGISTERED") == ""
    vid = _vol(3.0)
    aid = astra.algorithm.create({"type": "LAZY-SCALER", "ReconstructionDataId": vid,
                                  "options": {"Factor": -1.0}})
    astra.algorithm.run(aid, 1)
    assert np.array_equal(astra.data2d.get(vid), np.full((4, 5), -3.0, np.float32))
    astra.algorithm.delete(aid)
    astra.data2d.delete(vid)


def test_example_plugins_are_exposed():
    assert astra.plugins.SIRTPlugin.astra_name == "SIRT-PLUGIN"
    assert astra.plugins.CGLSPlugin.astra_name == "CGLS-PLUGIN"
    plugin.register(astra.plugins.SIRTPlugin)
    assert "'Relaxation'" in plugin.get_help("SIRT-PLUGIN")


def test_matlab_style_calls():
    """astra.m (python/astra/matlab.py): operation name first; 'iterate'
    is 'run' (shown with a host-only plugin)."""
    plugin.register(Scaler)
    vg = astra.create_vol_geom(3, 3)
    vid = astra.m.data2d("create", "-vol", vg, 2.0)
    aid = astra.m.algorithm("create", {"type": "SCALER-TEST", "ReconstructionDataId": vid,
                                       "options": {"Factor": 3.0}})
    astra.m.algorithm("iterate", aid, 2)
    assert np.array_equal(astra.m.data2d("get", vid), np.full((3, 3), 18.0, np.float32))
    assert astra.m.astra("use_cuda") in (True, False)
    with pytest.raises(astra.AstraError):
        astra.m.data2d("no_such_command", vid)
    mid = astra.m.matrix("create", __import__("scipy.sparse").sparse.eye(3, format="csr"))
    assert astra.m.matrix("get_size", mid) == (3, 3)
    astra.m.matrix("delete", mid)
    astra.m.algorithm("delete", aid)
    astra.m.data2d("delete", vid)


def test_sparse_matrix_store():
    """astra.matrix (python/astra/matrix_c.pyx:54-113): round trip, store
    within capacity, errors past it."""
    import scipy.sparse as ss
    rng = np.random.default_rng(0)
    m = ss.random(7, 11, density=0.3, random_state=1, format="csr", dtype=np.float32)
    mid = astra.matrix.create(m)
    assert astra.matrix.get_size(mid) == (7, 11)
    got = astra.matrix.get(mid)
    assert (got != m).nnz == 0 and got.dtype == np.float32
    smaller = ss.random(7, 11, density=0.1, random_state=2, format="csr", dtype=np.float32)
    astra.matrix.store(mid, smaller)
    assert (astra.matrix.get(mid) != smaller).nnz == 0
    with pytest.raises(astra.AstraError, match="too large"):
        astra.matrix.store(mid, ss.csr_matrix(rng.random((7, 11), dtype=np.float32)))
    with pytest.raises(astra.AstraError):
        astra.matrix.create(np.eye(3))
    astra.matrix.delete(mid)
    with pytest.raises(astra.AstraError):
        astra.matrix.get(mid)
