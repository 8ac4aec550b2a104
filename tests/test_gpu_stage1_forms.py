"""The u8 stage-1 kernels (8 DM rows in step through the L1, 16 rows sharing
an LDS window) against the oracle, bit for bit, over subband widths, DM
ranges, output lengths, inputs shorter than the sweep and time super-tiles."""
import dataclasses
import sys

import numpy as np
import pytest

from astroflow_amd import capi, synth
from oracle import oracle as O
from tests.helpers import make_spec, oracle_spec, small_band

pytestmark = pytest.mark.gpu

FORM_L1, FORM_LDS16 = 8, 16


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noise(rng, n, nc):
    return np.clip(np.round(rng.normal(128, 16, (n, nc))), 0, 255).astype(np.uint8)


def run_forms(x, spec, forms=(0, 16)):
    ref, rmask, *_ = O.run_pipeline(x, oracle_spec(spec), O.IQRM_GPU)
    seen = {}
    for dpb in forms:
        s = dataclasses.replace(spec, dm_per_block=dpb)
        plan = capi.Plan(s)
        form = plan.info.stage1_form
        got, masks = plan.run_host(x, want_masks=True)
        plan.close()
        say(f"  dm_per_block {dpb}: stage-1 form {form}")
        assert np.array_equal(got, ref), f"dm_per_block {dpb} (form {form}) differs from the oracle"
        if spec.use_iqrm:
            assert np.array_equal(masks, rmask)
        seen[dpb] = form
    return seen


@pytest.mark.parametrize("sbw,ndm0,dmhi,dmstep", [(128, 32, 3000.0, 0.5), (64, 16, 2000.0, 1.0),
                                                   (32, 32, 5000.0, 1.0), (256, 8, 500.0, 0.5),
                                                   (16, 8, 800.0, 2.0)])
def test_subband_forms_equal_oracle(gpu_lib, sbw, ndm0, dmhi, dmstep):
    say(f"case: subband u8 1024ch, sbw {sbw} ndm0 {ndm0} DM 0-{dmhi}/{dmstep}")
    rng = np.random.default_rng(sbw + ndm0)
    b = small_band(1024, 1000.0, 1500.0)
    ndm = int(dmhi / dmstep) + 1
    # input one sweep plus 3000 samples, a little short of the DM sweep at the
    # top DM (the zero padding past the input is read)
    fref = float(b.freq[-1])
    sweep = int(4148.808 * dmhi * (1 / 1000.0 ** 2 - 1 / fref ** 2) / float(b.tsamp))
    x = noise(rng, sweep + 2500, 1024)
    synth.inject_narrowband_rfi(x, rng, frac=0.05, amp=40, n_stuck=4, pmax=512)
    synth.inject_pulse(x, b.freq, b.f_top, float(b.tsamp), dmhi * 0.4, 900, 2e-3, 48.0)
    spec = make_spec(x, b, 0.0, dmstep, ndm, method="subband", sbw=sbw, ndm0=ndm0, nout=3000,
                     use_iqrm=True, iqrm_c0=0, iqrm_c1=1023, iqrm_mode=1,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=1023)
    seen = run_forms(x, spec)
    assert seen[0] == FORM_L1


def test_subband_4096ch_chunk_geometry(gpu_lib):
    # the 4096-channel chunk's subband geometry (32 subbands, DM 0-3000/0.5,
    # 32 fine DMs per coarse DM) on a short output
    say("case: subband u8 4096ch 1000-1500 MHz, DM 0-3000/0.5, 1700 outputs")
    rng = np.random.default_rng(40)
    b = small_band(4096, 1000.0, 1500.0)
    fref = float(b.freq[-1])
    sweep = int(4148.808 * 3000.0 * (1 / 1000.0 ** 2 - 1 / fref ** 2) / float(b.tsamp))
    x = noise(rng, sweep + 1800, 4096)
    synth.inject_pulse(x, b.freq, b.f_top, float(b.tsamp), 1200.0, 300, 2e-3, 48.0)
    spec = make_spec(x, b, 0.0, 0.5, 6001, method="subband", sbw=128, ndm0=32, nout=1700,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=4096)
    seen = run_forms(x, spec, forms=(0,))
    assert seen[0] == FORM_L1


@pytest.mark.parametrize("nch,dmhi,nout", [(256, 200.0, 5000), (1024, 60.0, 4097), (700, 31.0, 100)])
def test_direct_forms_equal_oracle(gpu_lib, nch, dmhi, nout):
    say(f"case: direct u8 {nch}ch DM 0-{dmhi}/1, nout {nout}")
    rng = np.random.default_rng(nch)
    b = small_band(nch, 1000.0, 1500.0)
    fref = float(b.freq[-1])
    sweep = int(4148.808 * dmhi * (1 / 1000.0 ** 2 - 1 / fref ** 2) / float(b.tsamp))
    x = noise(rng, nout + sweep - 50, nch)
    x[::7] = 255   # saturated rows: the packed 16-bit lanes at their largest
    spec = make_spec(x, b, 0.0, 1.0, int(dmhi) + 1, method="direct", nout=nout,
                     use_zero_dm=False)
    seen = run_forms(x, spec)
    assert seen[0] == FORM_L1


def test_subband_forms_super_tiles(gpu_lib):
    # output in several time super-tiles (a small intermediate budget)
    say("case: subband u8 512ch, 4 super-tiles")
    rng = np.random.default_rng(77)
    b = small_band(512, 1100.0, 1500.0)
    x = noise(rng, 20000, 512)
    spec = make_spec(x, b, 0.0, 1.0, 400, method="subband", sbw=64, ndm0=16,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=512, inter_budget=2 << 20)
    plan = capi.Plan(spec)
    assert plan.info.stage1_form == FORM_L1 and plan.info.ndm_nom > 0
    plan.close()
    run_forms(x, spec)
