"""The row filter through the C ABI (astra_amd_filter_create +
astra_amd_filter_rows) against the FP64 oracle's zero-padded spectral
filtering (cuda/3d/fdk.cu:351-398, cuda/2d/fft.cu:52-131), for every
transform size of the register (Stockham) path, 16 to 4096 points, and
the LDS paths above it: odd and even row counts, a spectrum that is not
the ramp (so an index error in the multiply shows), and both row kernels
for the real-spectrum case compared with each other."""
import ctypes
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SIZES = [5, 8, 9, 20, 33, 64, 100, 200, 384, 700, 768, 1000, 1536, 2048, 3000]


def _filter_rows(n_det, half, rows):
    import torch
    from astra import _core, _lib
    h = ctypes.c_void_p()
    half = np.ascontiguousarray(half, dtype=np.float32)
    _lib.call("astra_amd_filter_create", n_det, half.ctypes.data_as(ctypes.c_void_p), len(half),
              ctypes.byref(h))
    t = torch.from_numpy(np.ascontiguousarray(rows, dtype=np.float32)).cuda()
    _lib.call("astra_amd_filter_rows", h, _core.ptr(t), rows.shape[0], None)
    torch.cuda.synchronize()
    out = t.cpu().numpy()
    _lib.call("astra_amd_filter_destroy", h)
    return out


def _spectrum(n_det, seed):
    from oracle import fdk as OF
    n = OF.next_pow2(2 * n_det)
    rng = np.random.default_rng(seed)
    # ramp times a random smooth-ish positive profile: real and even (a half
    # spectrum), different at every bin
    return OF.ramp_filter(n_det) * (0.5 + rng.random(n // 2 + 1))


@pytest.mark.parametrize("n_det", SIZES)
def test_filter_rows_vs_oracle(n_det):
    from oracle import fdk as OF
    rng = np.random.default_rng(n_det)
    for n_rows in (1, 7, 64):
        rows = rng.random((n_rows, n_det), dtype=np.float32) - 0.3
        H = _spectrum(n_det, n_det + n_rows)
        got = _filter_rows(n_det, H, rows)
        ref = OF.filter_rows(rows, H.astype(np.float32))
        e = rel_l2(got, ref)
        assert e <= 1e-5, (n_det, n_rows, e)


@pytest.fixture
def lds_radix4():
    key = "ASTRA_AMD_FILTER_LDS_RADIX4"
    old = os.environ.get(key)
    os.environ[key] = "1"
    yield
    if old is None:
        os.environ.pop(key, None)
    else:
        os.environ[key] = old


@pytest.mark.parametrize("n_det", [20, 384, 768, 1536])
def test_filter_rows_lds_kernel_vs_oracle(n_det, lds_radix4):
    """The LDS radix-2^2 row kernel (kept for per-row complex spectra) on the
    real-spectrum case agrees with the oracle as well."""
    from oracle import fdk as OF
    rows = np.random.default_rng(7).random((9, n_det), dtype=np.float32)
    H = _spectrum(n_det, 3)
    got = _filter_rows(n_det, H, rows)
    assert rel_l2(got, OF.filter_rows(rows, H.astype(np.float32))) <= 1e-5


def test_filter_rows_many_rows_full_size():
    """720 x 768 rows of 768 bins at a time with the Ram-Lak spectrum (one
    detector row of the 512^3 FDK workload's projections, every angle):
    the oracle on a sample of rows, and linearity (filtering 2x equals 2x
    the filtered rows, bit for bit: power-of-two scale)."""
    import torch
    from oracle import fdk as OF
    n_det, n_rows = 768, 720 * 64
    rng = np.random.default_rng(11)
    rows = rng.random((n_rows, n_det), dtype=np.float32)
    H = OF.ramp_filter(n_det)
    got = _filter_rows(n_det, H, rows)
    pick = rng.choice(n_rows, 200, replace=False)
    assert rel_l2(got[pick], OF.filter_rows(rows[pick], H.astype(np.float32))) <= 1e-5
    got2 = _filter_rows(n_det, H, 2 * rows)
    assert np.array_equal(got2, 2 * got)
    del torch


@pytest.mark.parametrize("wg", ["128", "256", "512"])
@pytest.mark.parametrize("n_det", [20, 384, 768, 1536])
def test_filter_rows_workgroup_sizes_equal(n_det, wg, monkeypatch):
    """Every workgroup size of the register FFT (transforms per group)
    gives the default's floats."""
    half = _spectrum(n_det, n_det + 7)
    rows = np.random.default_rng(n_det).random((37, n_det), dtype=np.float32)
    ref = _filter_rows(n_det, half, rows)
    monkeypatch.setenv("ASTRA_AMD_FILTER_WG", wg)
    got = _filter_rows(n_det, half, rows)
    assert np.array_equal(got, ref)
