"""The GPU self-check and GPU information (python/astra/tests.py,
tests/python/unit/test_tests_module.py, test_astra_module.py) on the GPU."""
import pytest

pytestmark = pytest.mark.gpu


def test_self_check_on_gpu(capsys):
    import astra
    assert astra.use_cuda() is True
    assert isinstance(astra.get_gpu_info(), str) and "MB" in astra.get_gpu_info()
    capsys.readouterr()
    astra.test()
    out = capsys.readouterr().out.split("\n")[:-1]
    assert len(out) == 4
    assert out[2].endswith("Ok") and out[3].endswith("Ok")
