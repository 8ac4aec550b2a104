"""File readers of the C-ABI (af_fil_* for SIGPROC, af_psrfits_* for PSRFITS
search mode) and the module classes over them.  No GPU.

SIGPROC: every dtype and both band directions read back what was written
(flipped to ascending), for any row range.  PSRFITS: files written here with
their own FITS writer (primary HDU + SUBINT binary table, big-endian), read
back as pol 0 or the mean of pols 0 and 1 (floor for integers), ascending
band, with fch1 = OBSFREQ - OBSBW/2, foff = |CHAN_BW|, tsamp = TBIN and the
start MJD as python/astroflow/io/psrfits.py:32-76 derives them."""
import numpy as np
import pytest

from astroflow_amd import _astroflow_core as core
from astroflow_amd import capi
from tests.helpers import write_psrfits, write_sigproc


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32])
@pytest.mark.parametrize("descending", [False, True])
def test_sigproc_rows_read_back(tmp_path, dtype, descending):
    rng = np.random.default_rng(7)
    x = rng.integers(0, np.iinfo(dtype).max, (3001, 37), dtype=np.uint64).astype(dtype)
    p = tmp_path / "f.fil"
    if descending:
        write_sigproc(str(p), x[:, ::-1], fch1=1400.0 + 36 * 0.5, foff=-0.5, tsamp=2e-4)
    else:
        write_sigproc(str(p), x, fch1=1400.0, foff=0.5, tsamp=2e-4)
    r = capi.FileReader(str(p))
    assert (r.ndata, r.nchans, r.dtype) == (3001, 37, np.dtype(dtype))
    assert r.info.fch1 == pytest.approx(1400.0) and r.info.foff == pytest.approx(0.5)
    assert r.info.flipped == int(descending) and r.info.tsamp == pytest.approx(2e-4)
    assert np.array_equal(r.read(), x)
    assert np.array_equal(r.read(1234, 567), x[1234:1801])
    assert r.read(3001, 0).shape == (0, 37)
    with pytest.raises(ValueError):
        r.read(2990, 20)
    r.close()
    f = core.Filterbank(str(p))
    assert np.array_equal(f.data.reshape(3001, 37), x)


def test_sigproc_errors(tmp_path):
    with pytest.raises(RuntimeError, match="Can not open"):
        capi.FileReader(str(tmp_path / "missing.fil"))
    bad = tmp_path / "bad.fil"
    bad.write_bytes(b"\x05\x00\x00\x00HELLO" + b"\0" * 100)
    with pytest.raises(ValueError, match="not a SIGPROC"):
        capi.FileReader(str(bad))


def expected_psrfits(data, chan_bw):
    """pol 0, or floor((pol0 + pol1) / 2) / float mean, ascending band."""
    if data.shape[1] == 1:
        y = data[:, 0, :]
    elif data.dtype == np.float32:
        y = ((data[:, 0, :] + data[:, 1, :]) / np.float32(2)).astype(np.float32)
    else:
        y = ((data[:, 0, :].astype(np.uint64) + data[:, 1, :]) // 2).astype(data.dtype)
    return np.ascontiguousarray(y[:, ::-1] if chan_bw < 0 else y)


@pytest.mark.parametrize("dtype,npol,chan_bw", [(np.uint8, 2, -0.25), (np.uint8, 1, 0.25),
                                                (np.uint8, 4, 0.25), (np.uint16, 2, 0.5),
                                                (np.float32, 2, -1.0)])
def test_psrfits_pol_average_and_band(tmp_path, dtype, npol, chan_bw):
    rng = np.random.default_rng(3 + npol)
    nsamp, nchan, nsblk = 256 * 6, 48, 256
    if dtype == np.float32:
        data = rng.standard_normal((nsamp, npol, nchan)).astype(np.float32)
    else:
        data = rng.integers(0, np.iinfo(dtype).max, (nsamp, npol, nchan),
                            dtype=np.uint64).astype(dtype)
    p = tmp_path / "s.sf"
    obsfreq, obsbw = 1250.0, chan_bw * nchan
    write_psrfits(str(p), data, obsfreq, obsbw, chan_bw, 6.4e-5, nsblk)
    f = core.PsrFits(str(p))
    want = expected_psrfits(data, chan_bw)
    assert f.data.dtype == np.dtype(dtype) and f.data.shape == (nsamp, nchan)
    assert np.array_equal(f.data, want)
    fch1 = obsfreq - obsbw / 2
    if chan_bw < 0:
        fch1 -= (nchan - 1) * abs(chan_bw)
    assert f.fch1 == pytest.approx(fch1) and f.foff == pytest.approx(abs(chan_bw))
    assert f.tsamp == pytest.approx(6.4e-5) and f.npol == npol
    assert f.mjd == pytest.approx(60000 + 3600 / 86400 + 0.25 / 86400)
    h = f.header
    assert (h.nchans, h.ndata) == (nchan, nsamp)
    r = capi.FileReader(str(p), kind="psrfits")
    assert np.array_equal(r.read(300, 700), want[300:1000])   # spans sub-integrations
    r.close()


def test_psrfits_pol_mean_does_not_wrap(tmp_path):
    # 200 and 100 average to 150 (a uint8 sum would wrap to 44 and give 22)
    data = np.empty((64, 2, 8), np.uint8)
    data[:, 0, :], data[:, 1, :] = 200, 100
    p = tmp_path / "w.sf"
    write_psrfits(str(p), data, 1400.0, 8.0, 1.0, 1e-4, 32, extra_cols=False)
    assert np.all(core.PsrFits(str(p)).data == 150)


def test_psrfits_errors(tmp_path):
    p = tmp_path / "x.fits"
    p.write_bytes(b"SIMPLE  =                    T".ljust(80) + b"END".ljust(80) + b" " * 2720)
    with pytest.raises(ValueError, match="SUBINT"):
        core.PsrFits(str(p))
    q = tmp_path / "y.fits"
    q.write_bytes(b"not fits at all" * 300)
    with pytest.raises(ValueError):
        core.PsrFits(str(q))


def test_psrfits_inconsistent_table_is_rejected(tmp_path):
    data = np.zeros((64, 1, 8), np.uint8)
    p = tmp_path / "t.sf"
    write_psrfits(str(p), data, 1400.0, 8.0, 1.0, 1e-4, 32, extra_cols=False)
    raw = bytearray(p.read_bytes())
    # NAXIS1 (the row size) smaller than the DATA column
    i = raw.index(b"NAXIS1  =")
    raw[i:i + 80] = b"NAXIS1  =                  100".ljust(80)
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="past the SUBINT row"):
        core.PsrFits(str(p))
