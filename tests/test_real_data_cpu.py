"""The oracle on real data: the reference's fixture tests/FRB180417.fil.
Dedispersing it over DM 0-1000 (both methods, IQRM + zero-DM) puts the
largest peak S/N at DM 475, next to the burst's published DM of 474.8, and
well above the DM-0 row."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import real_data as R


@pytest.mark.parametrize("method", ["direct", "subband"])
def test_oracle_recovers_the_burst_at_its_published_dm(method):
    x, freq, tsamp, _ = R.load()
    plane, masks, t0s, t1s, _ = O.run_pipeline(x, R.oracle_spec(freq, tsamp, method))
    assert plane.shape == (1001, 5120)
    m, snr, t = R.best_dm(plane)
    assert abs(m - R.FRB_DM) <= 2.5, (m, snr)
    assert snr > 7.0 and abs(t - 1602) <= 3
    assert R.best_dm(plane[:1])[1] < 5.0          # no burst at DM 0
    assert masks.shape == (1, 336) and 0 < masks.sum() < 34
