"""`_dedispered_fil(..., target=0)`: the host-core path of the drop-in
(cpucal::dedispered_fil_omp, include/cpucal.hpp:75-201).  No GPU.

The images are compared with the oracle: its FP64 delay table at the native
rate and its restatement of the cpucal loop nest (oracle/af_cpucal.c, itself
pinned to the direct method in test_oracle.py), per slice of
round(t_sample / tsamp) samples, imaged by the module's _image_slice (pinned
to closed forms in test_bindings_cpu.py).  u16 files are checked against a
brute-force sum on a tiny input."""
import numpy as np
import pytest

from astroflow_amd import _astroflow_core as core
from oracle import oracle as O
from tests.helpers import write_sigproc

REF_FIL = "tests/golden/FRB180417.fil"


def oracle_images(x, fch1, foff, tsamp, dm_low, dm_high, fs, fe, dm_step, ds, t_sample):
    """x: [nsamp][nchan] ascending band, as the reader returns it."""
    nsamp, nch = x.shape
    freq = O.freq_table_f32(fch1, foff, nch)
    cs, ce = O.chan_range(freq, fs, fe)
    ndm = int((np.float32(dm_high) - np.float32(dm_low)) / np.float32(dm_step)) + 1
    dms = float(np.float32(dm_low)) + np.arange(ndm, dtype=np.float64) * float(np.float32(dm_step))
    # delays with the header's float32 sample time; slices from the file's
    d = O.delay_table(freq, cs, ce, freq[ce], float(np.float32(tsamp)), 1, dms)
    sps = int(round(float(np.float32(t_sample)) / tsamp))
    out = []
    for start in range(0, nsamp, sps):
        n = min(sps, nsamp - start)
        plane = O.cpucal_direct_u8(x, cs, ce, d, start, n, ds)
        out.append(core._image_slice(np.ascontiguousarray(plane.astype(np.float32)), 512))
    return out


def noisy_with_pulse(nsamp, nch, fch1, foff, tsamp, dm, seed):
    rng = np.random.default_rng(seed)
    x = np.clip(np.rint(rng.normal(100, 12, (nsamp, nch))), 0, 255).astype(np.uint8)
    f = fch1 + foff * np.arange(nch)
    d = np.rint(4148.808 * dm * (1 / f**2 - 1 / f.max()**2) / tsamp).astype(int)
    for c in range(nch):
        t = nsamp // 3 + d[c]
        if t < nsamp:
            x[t, c] = 250
    return x


@pytest.mark.parametrize("ds,t_sample,descending", [(1, 0.2, False), (2, 0.15, True),
                                                    (3, 0.5, False)])
def test_u8_file_equals_the_oracle(tmp_path, ds, t_sample, descending):
    nsamp, nch, fch1, foff, tsamp = 6000, 96, 1200.0, 3.0, 1e-4
    x = noisy_with_pulse(nsamp, nch, fch1, foff, tsamp, 120.0, 11 + ds)
    p = tmp_path / "beam.fil"
    if descending:
        write_sigproc(str(p), x[:, ::-1], fch1=fch1 + foff * (nch - 1), foff=-foff, tsamp=tsamp)
    else:
        write_sigproc(str(p), x, fch1=fch1, foff=foff, tsamp=tsamp)
    args = (0.0, 200.0, 1200.0, 1485.0, 2.0)
    res = core._dedispered_fil(str(p), *args, ds, t_sample, 0)
    ref = oracle_images(x, fch1, foff, tsamp, *args, ds, t_sample)
    assert len(res.dm_times) == len(ref) > 1
    assert res.shape == [512, 512, 3] and res.tsample == pytest.approx(t_sample)
    for k, (a, b) in enumerate(zip(res.dm_times, ref)):
        assert np.array_equal(a, b), k


def test_reference_fixture_file_equals_the_oracle():
    f = core.Filterbank(REF_FIL)
    x = np.ascontiguousarray(f.data.reshape(f.ndata, f.nchans))
    h = f.header
    fe = float(np.float32(h.fch1) + np.float32(h.foff) * (f.nchans - 1))
    args = (300.0, 700.0, float(h.fch1), fe, 1.0)
    res = core._dedispered_fil(REF_FIL, *args, 1, 0.25, 0)
    ref = oracle_images(x, h.fch1, h.foff, f.tsamp, *args, 1, 0.25)
    assert len(res.dm_times) == len(ref) > 0
    for a, b in zip(res.dm_times, ref):
        assert np.array_equal(a, b)


def test_u16_file_equals_brute_force(tmp_path):
    rng = np.random.default_rng(2)
    nsamp, nch, fch1, foff, tsamp = 700, 12, 1300.0, 10.0, 5e-4
    x = rng.integers(0, 60000, (nsamp, nch)).astype(np.uint16)
    p = tmp_path / "b16.fil"
    write_sigproc(str(p), x, fch1=fch1, foff=foff, tsamp=tsamp)
    dm_low, dm_high, dm_step, ds, t_sample = 0.0, 30.0, 3.0, 2, 0.2
    res = core._dedispered_fil(str(p), dm_low, dm_high, 1300.0, 1410.0, dm_step, ds, t_sample, 0)
    freq = O.freq_table_f32(fch1, foff, nch)
    cs, ce = O.chan_range(freq, 1300.0, 1410.0)
    sps = int(round(t_sample / tsamp))
    ndm = 11
    for k, start in enumerate(range(0, nsamp, sps)):
        n = min(sps, nsamp - start)
        down = -(-n // ds)
        plane = np.zeros((ndm, down), np.uint64)
        for i in range(ndm):
            dm = dm_low + i * dm_step
            for c in range(cs, ce):
                dly = 4148.808 * dm * (1 / freq[c] ** 2 - 1 / freq[ce] ** 2) / float(np.float32(tsamp))
                dd = int(np.floor(dly + 0.5))
                for ti in range(down):
                    t = start + ti * ds + dd
                    if t < nsamp:
                        plane[i, ti] += int(x[t, c])
        want = core._image_slice(plane.astype(np.float32), 512)
        assert np.array_equal(res.dm_times[k], want), k


def test_argument_errors(tmp_path):
    x = np.zeros((400, 8), np.uint8)
    p = tmp_path / "z.fil"
    write_sigproc(str(p), x, fch1=1000.0, foff=1.0, tsamp=1e-3)
    with pytest.raises(ValueError, match="out of filterbank range"):
        core._dedispered_fil(str(p), 0.0, 10.0, 990.0, 1007.0, 1.0, 1, 0.1, 0)
    with pytest.raises(ValueError, match="Invalid chan"):
        core._dedispered_fil(str(p), 0.0, 10.0, 1003.0, 1003.0, 1.0, 1, 0.1, 0)
    with pytest.raises(ValueError, match="Invalid DM"):
        core._dedispered_fil(str(p), 10.0, 0.0, 1000.0, 1007.0, 1.0, 1, 0.1, 0)
    with pytest.raises(ValueError, match="exceeds total observation"):
        core._dedispered_fil(str(p), 0.0, 10.0, 1000.0, 1007.0, 1.0, 1, 1.0, 0)
    with pytest.raises(ValueError, match="target"):
        core._dedispered_fil(str(p), 0.0, 10.0, 1000.0, 1007.0, 1.0, 1, 0.1, 2)
