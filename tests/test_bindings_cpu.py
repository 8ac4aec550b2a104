"""The `_astroflow_core` module without a GPU: names, constructors and
defaults of the reference's classes, the filterbank reader, and the DM-time
imaging against closed forms (identity resize, block-average downsizing,
min-max scaling and the viridis table)."""
import os

import numpy as np
import pytest

from astroflow_amd import capi
from tests.helpers import write_sigproc

core = pytest.importorskip("astroflow_amd._astroflow_core")


def test_reference_names_exist():
    for name in ("_dedisperse_spec", "_dedisperse_spec_raw", "_dedispered_fil", "Header",
                 "IQRMconfig", "RFIconfig", "DedispersedDataUint8", "DedisperedData",
                 "Filterbank"):
        assert hasattr(core, name), name


def test_config_constructors_and_defaults():
    # defaults of include/rfi.h:14-16,43
    q = core.IQRMconfig()
    assert (q.mode, q.nsigma, q.geofactor, q.win_sec, q.hop_sec, q.include_tail) == \
        (0, 3.0, 1.5, 5.0, 2.5, True)
    assert q.radius_frac == pytest.approx(0.1)
    r = core.RFIconfig()
    assert (r.use_mask, r.use_zero_dm, r.use_iqrm) == (False, False, False)
    # positional order of the Python subclasses (config/rficonfig.py:20,47)
    q2 = core.IQRMconfig(1, 0.25, 4.0, 2.0, 0.0, 0.5, False)
    assert (q2.mode, q2.nsigma, q2.geofactor, q2.hop_sec, q2.include_tail) == (1, 4.0, 2.0, 0.5, False)
    r2 = core.RFIconfig(True, True, False, q2)
    assert (r2.use_mask, r2.use_zero_dm, r2.use_iqrm, r2.iqrm_cfg.mode) == (True, True, False, 1)

    class Obj:
        pass
    o = Obj()
    o.use_mask, o.use_zero_dm, o.use_iqrm = False, True, True
    o.iqrm_cfg = Obj()
    for k, v in dict(mode=1, radius_frac=0.05, nsigma=3.0, geofactor=1.5, win_sec=0.0,
                     hop_sec=0.0, include_tail=True).items():
        setattr(o.iqrm_cfg, k, v)
    r3 = core.RFIconfig(o)
    assert r3.use_iqrm and r3.use_zero_dm and r3.iqrm_cfg.mode == 1
    h = core.Header(60000.0, 1e-4, 1000.0, 0.5, 8, 1024, 5000, "x.fil")
    assert (h.nchans, h.ndata, h.nbits, h.filename) == (1024, 5000, 8, "x.fil")
    h.__class__  # noqa: B018


@pytest.mark.skipif(capi.device_count() > 0, reason="a GPU is visible")
def test_call_without_gpu_raises_runtime_error():
    h = core.Header(0.0, 1e-4, 1000.0, 0.5, 8, 64, 100, "x")
    with pytest.raises(RuntimeError, match="No HIP devices"):
        core._dedisperse_spec_raw(np.zeros((100, 64), np.uint8), h, 0.0, 10.0, 1000.0, 1031.0)


def test_filterbank_reader_flips_descending_band(tmp_path):
    rng = np.random.default_rng(0)
    x = rng.integers(0, 255, (300, 48), dtype=np.uint8)
    p = tmp_path / "t.fil"
    write_sigproc(str(p), x, fch1=1500.0, foff=-2.0, tsamp=6.4e-5)
    f = core.Filterbank(str(p))
    assert (f.nchans, f.nbits, f.ndata, f.nifs) == (48, 8, 300, 1)
    assert f.foff == 2.0 and f.fch1 == 1500.0 - 47 * 2.0
    assert np.array_equal(f.data[:, 0, :], x[:, ::-1])
    h = f.header
    assert (h.nchans, h.ndata) == (48, 300) and h.fch1 == np.float32(1406.0)


def test_filterbank_reader_16bit_ascending(tmp_path):
    x = np.arange(20 * 5, dtype=np.uint16).reshape(20, 5) * 300
    p = tmp_path / "u16.fil"
    write_sigproc(str(p), x, fch1=1000.0, foff=1.0, tsamp=1e-3)
    f = core.Filterbank(str(p))
    assert f.nbits == 16 and np.array_equal(f.data[:, 0, :], x)


# the reference's own test fixture (tests/FRB180417.fil), kept unchanged under tests/golden/
REF_FIL = os.path.join(os.path.dirname(__file__), "golden", "FRB180417.fil")


def test_reads_the_reference_fixture_file():
    # tests/FRB180417.fil of the reference: 336 channels, 8-bit, 5120 samples
    f = core.Filterbank(REF_FIL)
    assert (f.nchans, f.nbits, f.nifs) == (336, 8, 1)
    assert f.ndata == 5120
    assert f.foff > 0
    d = f.data
    assert d.shape == (5120, 1, 336) and d.dtype == np.uint8


def _viridis_bgr():
    from matplotlib import colormaps
    rgb = np.asarray(colormaps["viridis"].colors)
    return np.rint(rgb * 255).astype(np.uint8)[:, ::-1]


def test_identity_resize_minmax_and_viridis_table():
    # 512 x 512 keeps its size (the 8-tap kernel at integer positions is the
    # identity); values 0..255 with min 0 and max 255 map to themselves
    g = (np.arange(512 * 512) % 256).reshape(512, 512).astype(np.float32)
    img = core._image_slice(g, 512)
    lut = _viridis_bgr()
    assert img.shape == (512, 512, 3)
    assert np.array_equal(img, lut[g.astype(np.int64)])


def test_area_downsizing_is_block_average():
    # 1024 x 1536 -> 512 x 512: every output averages a 2 x 3 block; blocks
    # are constant so the average is the block value
    rng = np.random.default_rng(1)
    blocks = rng.integers(0, 256, (512, 512)).astype(np.float32)
    blocks[0, 0], blocks[0, 1] = 0, 255
    src = np.repeat(np.repeat(blocks, 2, axis=0), 3, axis=1)
    img = core._image_slice(np.ascontiguousarray(src), 512)
    assert np.array_equal(img, _viridis_bgr()[blocks.astype(np.int64)])


def test_flat_slice_maps_to_the_first_colour():
    # max == min: NORM_MINMAX scales by 0 (a flat slice at the identity size,
    # where resizing is exact)
    img = core._image_slice(np.full((512, 512), 7.0, np.float32), 512)
    assert np.all(img == _viridis_bgr()[0])


def _rfi_iqrm():
    return core.RFIconfig(False, True, True, core.IQRMconfig(1, 0.1, 3.0, 1.5, 0.0, 0.0, True))


def test_derived_description_follows_the_reference():
    # 64 channels 1000.0 + 0.5 i MHz: freq 1004-1020 maps to channels 8..40
    # (src/core/dedispered.cu:1017-1022); IQRM and zero-DM take [cs, ce)
    h = core.Header(0.0, 1e-4, 1000.0, 0.5, 8, 64, 1000, "x")
    d = core._describe_spec(1000, h, 0.0, 10.0, 1004.0, 1020.0, 0.5, 2, "", _rfi_iqrm())
    assert (d["chan_start"], d["chan_end"]) == (8, 40)
    assert (d["dd_c0"], d["dd_c1"]) == (8, 41)
    assert (d["use_iqrm"], d["iqrm_c0"], d["iqrm_c1"]) == (1, 8, 40)
    assert (d["use_zero_dm"], d["zdm_c0"], d["zdm_c1"]) == (1, 8, 40)
    assert d["dm_steps"] == 21 and d["down_ndata"] == 500
    assert d["f_ref_mhz"] == pytest.approx(1020.0)


def test_single_channel_range_skips_iqrm():
    # freq_start == freq_end == the top channel: cs == ce, so IQRM's channel
    # range [cs, ce) is empty and the reference computes no window masks
    # (src/core/iqrm.cu:363-364); the call goes ahead without IQRM
    h = core.Header(0.0, 1e-4, 1000.0, 0.5, 8, 64, 1000, "x")
    top = 1000.0 + 0.5 * 63
    d = core._describe_spec(1000, h, 0.0, 10.0, top, top, 1.0, 1, "", _rfi_iqrm())
    assert d["chan_start"] == d["chan_end"] == 63
    assert (d["dd_c0"], d["dd_c1"]) == (63, 64)
    assert d["use_iqrm"] == 0
    assert d["zdm_c0"] == d["zdm_c1"] == 63
