"""Toolbox-level functions (tests/python/unit/test_astra_module.py,
test_tests_module.py, test_algorithm_module.py, test_projector_module.py):
credits, feature flags, delete / info by id across object kinds, the
object listings, clipCircle, and the self-check without a GPU."""
import numpy as np
import pytest

import astra


def test_credits_and_flags(capsys):
    astra.astra.credits()
    assert len(capsys.readouterr().out) > 0
    for f in ("cuda", "projectors_scaled_as_line_integrals", "fan_cone_BP_density_weighting_by_default",
              "unpadded_GPULink"):
        assert astra.astra.has_feature(f) is True
    assert astra.astra.has_feature("random_string") is False
    assert isinstance(astra.astra.use_cuda(), bool)


def test_delete_and_info_by_id(capsys):
    i = astra.data3d.create("-vol", astra.create_vol_geom(10, 10, 10))
    j = astra.data2d.create("-sino", astra.create_proj_geom("parallel", 1.0, 8, [0.0, 1.0]))
    capsys.readouterr()
    astra.astra.info([i, j])
    lines = capsys.readouterr().out.split("\n")[:-1]
    assert lines == ["data3d\t%d \tv     10x10x10 volume data \t" % i,
                     "data2d\t%d \tv     8x2 sinogram data \t" % j]
    astra.astra.delete(i)
    astra.astra.info(i)
    assert capsys.readouterr().out == ""
    with pytest.raises(astra.log.AstraError):
        astra.data3d.get(i)
    astra.astra.delete([j, 10 ** 6])     # unknown ids are ignored
    with pytest.raises(astra.log.AstraError):
        astra.data2d.get(j)


def test_projector_and_algorithm_listings(capsys):
    pg = astra.create_proj_geom("parallel3d", 1.0, 1.0, 20, 20, np.linspace(0, 1, 30))
    vg = astra.create_vol_geom(10, 10, 10)
    astra.projector3d.clear()
    astra.algorithm.clear()
    pid = astra.create_projector("cuda3d", pg, vg)
    p2 = astra.create_projector("cuda", astra.create_proj_geom("parallel", 1.0, 8, [0.0]),
                                astra.create_vol_geom(4, 4))
    assert astra.projector.is_cuda(p2) and astra.projector3d.is_cuda(pid)
    cfg = astra.astra_dict("FP3D_CUDA")
    cfg["VolumeDataId"] = astra.data3d.create("-vol", vg)
    cfg["ProjectionDataId"] = astra.data3d.create("-sino", pg)
    aid = astra.algorithm.create(cfg)
    capsys.readouterr()
    for mod, want in ((astra.projector3d, "%d \tv     " % pid), (astra.projector, "%d \tv     cuda" % p2),
                      (astra.algorithm, "%d \tv     FP3D_CUDA" % aid)):
        mod.info()
        out = capsys.readouterr().out
        assert len(out.split("\n")) - 5 == 1 and out.split("\n")[2] == want
    astra.algorithm.delete(aid)
    astra.algorithm.info()
    assert len(capsys.readouterr().out.split("\n")) - 5 == 0
    with pytest.raises(astra.log.AstraError):
        astra.algorithm.run(aid)
    astra.projector3d.delete(pid)
    astra.projector.delete(p2)
    astra.data3d.delete([cfg["VolumeDataId"], cfg["ProjectionDataId"]])


@pytest.mark.parametrize("n", [1, 6, 7, 32])
def test_clip_circle(n):
    img = np.random.default_rng(n).random((n, n)) + 1.0
    want = img.copy()
    zeroed = 0
    mid = (n - 1) / 2.0
    for i in range(n):
        for j in range(n):
            if (i - mid) ** 2 + (j - mid) ** 2 > n * n / 4.0:
                want[i, j] = 0
                zeroed += 1
    assert astra.clipCircle(img) == zeroed
    assert np.array_equal(img, want)


def test_self_check_without_gpu(capsys):
    astra.test_noCUDA()
    assert len(capsys.readouterr().out.split("\n")) - 1 == 2


def test_log_module(tmp_path, capsys):
    """astra.log (python/astra/log.py): levels, screen and file outputs,
    formats with the reference's substitutions, enable / disable."""
    log = astra.log
    log.setOutputScreen(log.STDOUT, log.WARN)
    log.setFormatScreen("%l|%m|%%\n")
    log.info("hidden")
    log.warn("shown")
    assert capsys.readouterr().out == "WARN|shown|%\n"
    path = tmp_path / "astra.log"
    log.setOutputFile(str(path), log.DEBUG)
    log.setFormatFile("%l %f:%n %m\n")
    log.debug("to file only")
    log.disableScreen()
    log.error("file and no screen")
    log.enableScreen()
    log.disableFile()
    log.error("screen only")
    text = path.read_text()
    assert "DEBUG " in text and "to file only" in text and "file and no screen" in text
    assert "screen only" not in text and "test_astra_module.py:" in text
    out = capsys.readouterr().out
    assert out == "ERROR|screen only|%\n"
    log.disable()
    log.error("nowhere")
    assert capsys.readouterr().out == ""
    log.enable()
    log.setOutputScreen(log.STDERR, log.ERROR)
    log.setFormatScreen("%d %t %f(%n): %l: %m\n")
    e = astra.AstraError("plain", append_log=False)
    assert str(e) == "plain"
