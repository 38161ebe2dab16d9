"""The C ABI: libastra_amd.so loads and exports every entry point that
include/astra_amd.h declares; argument errors come back as status codes with
a message (host-side validation only, no kernel launch)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "astra_amd.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(astra_amd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_hot_path():
    names = declared()
    for must in ("astra_amd_fp3d", "astra_amd_bp3d", "astra_amd_fp2d", "astra_amd_bp2d",
                 "astra_amd_fdk", "astra_amd_filter_rows", "astra_amd_algo_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from astra import _lib
    lib = _lib.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes signature table covers them all
    assert set(declared()) <= set(_lib.SIGNATURES) | set(_lib.STRINGS)
    assert "gfx950" in _lib.version()


def test_argument_errors_are_reported():
    from astra import _lib
    lib = _lib.load()
    out = ctypes.c_void_p()
    st = lib.astra_amd_geom3d_create(99, 4, 4, 4, 4, 4, 4, None, None, ctypes.byref(out))
    assert st == _lib.ERR_ARG
    assert b"kind" in lib.astra_amd_last_error()
    st = lib.astra_amd_geom2d_create(_lib.GEOM_PAR, 0, 4, 4, 4, None, ctypes.byref(out))
    assert st == _lib.ERR_ARG
    st = lib.astra_amd_fp3d(None, None, None, None, 0, None)
    assert st == _lib.ERR_ARG
    assert b"NULL" in lib.astra_amd_last_error()
    with pytest.raises(Exception):
        _lib.call("astra_amd_algo_run", None, None, None, 1, None)


def test_filter_spectra_match_the_oracle():
    """astra_amd_filter_spectrum (host-only) for every window against the
    FP64 oracle's genFilter restatement, float32 rounding apart."""
    import numpy as np
    from astra import _lib
    from oracle import fdk as OF
    for n_det, d, param in ((50, 1.0, -1.0), (97, 0.7, -1.0), (64, 1.0, 0.3)):
        n = OF.next_pow2(2 * n_det)
        nb = n // 2 + 1
        for t, name in enumerate(OF.FILTER_TYPES):
            out = np.zeros(nb, dtype=np.float32)
            _lib.call("astra_amd_filter_spectrum", t, n_det, d, param,
                      out.ctypes.data_as(ctypes.c_void_p), nb)
            ref = OF.windowed_filter(name, n_det, d, param)
            np.testing.assert_allclose(out, ref, rtol=1e-6, atol=1e-7 * np.abs(ref).max(),
                                       err_msg=name)
    with pytest.raises(Exception):
        _lib.call("astra_amd_filter_spectrum", 99, 10, 1.0, -1.0, None, 0)


def test_abort_hook_set_and_cleared():
    """astra_amd_set_abort_hook is host-only: a hook and NULL are accepted;
    the Python layer installs its hook only for the duration of a run."""
    import signal
    from astra import _lib, algorithm
    _lib.call("astra_amd_set_abort_hook", algorithm._native_hook, None)
    _lib.call("astra_amd_set_abort_hook", algorithm._no_hook, None)
    before = signal.getsignal(signal.SIGINT)
    with algorithm._interruptible():
        assert signal.getsignal(signal.SIGINT) is not before
        signal.raise_signal(signal.SIGINT)  # sets the stop flag, raises nothing
        assert algorithm._hook(None) == 1
    assert signal.getsignal(signal.SIGINT) is before


def test_ignored_sigint_is_left_alone():
    import signal
    from astra import algorithm
    prev = signal.signal(signal.SIGINT, signal.SIG_IGN)
    try:
        with algorithm._interruptible():
            assert signal.getsignal(signal.SIGINT) is signal.SIG_IGN
    finally:
        signal.signal(signal.SIGINT, prev)
