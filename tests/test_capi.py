"""The C ABI library loads and exports every entry point include/mxr.h declares (no GPU)."""
import ctypes
import os
import re

import pytest

import mxr
import oracle_lib as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "mxr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mxr_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(mxr.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n


def test_create_without_device_is_an_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = mxr.lib().mxr_create(ctypes.byref(h), 0, 64)
    assert rc == mxr.MXR_ERR_NO_DEVICE
    assert mxr.lib().mxr_strerror(rc).decode() == "no usable gfx950 device"


def test_create_rejects_unsupported_dim():
    h = ctypes.c_void_p()
    assert mxr.lib().mxr_create(ctypes.byref(h), 0, 100) == mxr.MXR_ERR_INVALID_ARG
    assert b"not supported" in mxr.lib().mxr_last_error()


@pytest.mark.parametrize("N,box,n", [(64, 67.8, 44), (256, 200.0, 44), (512, 400.0, 44), (96, 100.0, 46),
                                     (128, 125.0, 42), (160, 150.0, 42)])
def test_radius_schedule_matches_oracle(N, box, n):
    p = mxr.hii_params(N, box)
    r = mxr.radii(p)
    assert len(r) == n
    assert r == O.radii(p)
    assert r[0] == 50.0 or r[0] == pytest.approx(0.620350491 * box)


@pytest.mark.parametrize("N,box,h", [(16, 20.0, 0.678), (64, 67.8, 0.678), (512, 400.0, 0.678), (160, 150.0, 0.7)])
def test_ps_length_matches_oracle(N, box, h):
    assert mxr.ps_length(N, box, h) == O.ps_length(N, box, h)


def test_rtom_matches_oracle():
    p = mxr.hii_params(64, 67.8)
    for R in (0.5, 3.0, 42.0):
        assert mxr.lib().mxr_RtoM(ctypes.byref(p), R) == O.lib().oracle_RtoM(ctypes.byref(p), R)
    p.ReionRtoMFilterType = 1
    assert mxr.lib().mxr_RtoM(ctypes.byref(p), 2.0) == O.lib().oracle_RtoM(ctypes.byref(p), 2.0)


def test_null_arguments_are_errors():
    L = mxr.lib()
    assert L.mxr_find_hii_bubbles(None, None, None, None) == mxr.MXR_ERR_INVALID_ARG
    assert L.mxr_set_path(None, 0) == mxr.MXR_ERR_INVALID_ARG
    t = mxr.Tuning()
    assert L.mxr_get_tuning(None, ctypes.byref(t)) == mxr.MXR_ERR_INVALID_ARG
    assert L.mxr_set_tuning(None, ctypes.byref(t)) == mxr.MXR_ERR_INVALID_ARG
    assert L.mxr_profile_enable(None, 1) == mxr.MXR_ERR_INVALID_ARG
    assert L.mxr_power_spectrum(None, None, 0, 400.0, 0.678, 21, None, None, None) == mxr.MXR_ERR_INVALID_ARG
    assert L.mxr_ps_counts(None, 400.0, 0.678, 21, None) == mxr.MXR_ERR_INVALID_ARG
    assert L.mxr_ps_length(0, 400.0, 0.678) == mxr.MXR_ERR_INVALID_ARG
    assert L.mxr_ps_length(512, -1.0, 0.678) == mxr.MXR_ERR_INVALID_ARG


def test_supported_dims_match_header():
    """The dimension list include/mxr.h states at mxr_supported_dims is the build's list."""
    import re

    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mxr.h")).read()
    block = hdr[hdr.index("The grid dimensions this build compiles"):hdr.index("int mxr_supported_dims")]
    listed = [int(x) for x in re.search(r"\*\s+([0-9, ]+)\n", block).group(1).replace(" ", "").split(",") if x]
    assert mxr.supported_dims() == listed
    assert listed == sorted(listed)
