"""The CPU baseline's backend: the oracle's FP32 restatement with its transforms on FFTW3's float
API (libmkl_rt's export) gives the restatement's own results to the parity bounds."""
import numpy as np
import pytest

import mxr
import oracle_lib as O
from conftest import realisation
from parity import check, compare


@pytest.mark.skipif(not O.fftw_available(), reason="no library exporting FFTW3's float API")
@pytest.mark.parametrize("N,box", [(32, 40.0), (64, 67.8)])
def test_fftw_backend_matches_restatement(N, box):
    inp = realisation(N, box, 1)
    p = mxr.hii_params(N, box, filter_type=1)
    sa, sb = mxr.HiiState.host(N), mxr.HiiState.host(N)
    ga, margin = O.find_hii_bubbles(p, inp, sa, margin=True)
    gb, secs = O.find_hii_bubbles_fftw(p, inp, sb)
    st = compare(N, {k: np.asarray(v) for k, v in sb.arrays().items()},
                 {k: np.asarray(v) for k, v in sa.arrays().items()}, gb, ga, margin)
    check(st)
    assert gb.n_radii == ga.n_radii and secs > 0
    assert 0.05 < ga.volume_weighted_global_xH < 0.95
