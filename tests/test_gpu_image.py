"""DM-time imaging on the device (af_image_slices, af_image.hip) against the
host imaging module (af_image.h, exposed as _image_slice) and against closed
forms.

  - random planes (float32, u32, u64), several slices with a narrower last
    one, in the three resize regimes (area on both axes, Lanczos on both,
    upsizing): every image equals _image_slice of the same slice bit for bit;
  - closed forms computed on the device: the identity size maps values
    0..255 with min 0 / max 255 straight through the colour table; a 2 x 3
    area shrink of constant blocks is the block value; a flat slice maps to
    the first colour;
  - the colour table is matplotlib's viridis (rounded, BGR).
OpenCV is not installed, so agreement with cv::resize / cv::normalize /
cv::applyColorMap themselves is not checked here (parity with OpenCV is
unpinned); the closed forms are what those definitions fix."""
import sys

import numpy as np
import pytest

from astroflow_amd import capi

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def viridis_bgr():
    from matplotlib import colormaps
    rgb = np.asarray(colormaps["viridis"].colors)
    return np.rint(rgb * 255).astype(np.uint8)[:, ::-1]


CODES = {np.float32: capi.AF_F32, np.uint32: capi.AF_U32, np.uint64: capi.AF_U64}


def gpu_images(plane: np.ndarray, sps: int, size: int = 512, cols: int = None):
    rows, ld = plane.shape
    cols = ld if cols is None else cols
    nsl = -(-cols // sps)
    d = capi.DeviceBuffer(plane.nbytes)
    d.upload(np.ascontiguousarray(plane))
    o = capi.DeviceBuffer(nsl * size * size * 3)
    try:
        capi.check(capi.lib().af_image_slices(capi.C.c_void_p(d.ptr), CODES[plane.dtype.type],
                                              rows, cols, ld, sps, size,
                                              capi.C.c_void_p(o.ptr), None),
                   "af_image_slices")
        return o.download(np.empty((nsl, size, size, 3), np.uint8))
    finally:
        d.free()
        o.free()


def host_images(plane: np.ndarray, sps: int, size: int = 512, cols: int = None):
    from astroflow_amd import _astroflow_core as core
    cols = plane.shape[1] if cols is None else cols
    out = []
    for c0 in range(0, cols, sps):
        sl = np.ascontiguousarray(plane[:, c0:min(cols, c0 + sps)].astype(np.float32))
        out.append(core._image_slice(sl, size))
    return np.stack(out)


@pytest.mark.parametrize("rows,cols,sps,dtype", [
    (700, 3000, 1100, np.float32),   # area on both axes, last slice 800 wide
    (300, 2000, 700, np.float32),    # Lanczos on both axes (rows < 512), last 600
    (100, 450, 200, np.float32),     # upsizing on both axes, last 50
    (1001, 2600, 1300, np.uint32),   # u32 plane (u8/u16 samples), two full slices
    (640, 1500, 1024, np.uint64),    # u64 plane (u32 samples), last 476 (Lanczos)
])
def test_device_images_equal_host_imaging(gpu_lib, rows, cols, sps, dtype):
    say(f"case: imaging {rows} x {cols}, {sps} columns per slice, {np.dtype(dtype).name}")
    rng = np.random.default_rng(rows + cols)
    if dtype == np.float32:
        plane = rng.standard_normal((rows, cols)).astype(np.float32) * 50.0
        plane[rows // 3, cols // 2] = 900.0   # a bright point, so slices differ in range
    else:
        plane = rng.integers(0, 1 << 20, (rows, cols)).astype(dtype)
        if dtype == np.uint64:
            plane += np.uint64(1) << np.uint64(40)   # above 2^32: the u64 conversion
    got = gpu_images(plane, sps)
    ref = host_images(plane, sps)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_columns_beyond_cols_are_not_imaged(gpu_lib):
    say("case: imaging with ld > cols (a plane row pitch wider than its valid columns)")
    rng = np.random.default_rng(9)
    plane = rng.standard_normal((520, 1400)).astype(np.float32)
    got = gpu_images(plane, 600, cols=1300)
    assert got.shape[0] == 3
    assert np.array_equal(got, host_images(plane, 600, cols=1300))


def test_identity_size_is_the_colour_table(gpu_lib):
    say("case: imaging closed form, identity size")
    g = (np.arange(512 * 512) % 256).reshape(512, 512).astype(np.float32)
    got = gpu_images(g, 512)
    assert np.array_equal(got[0], viridis_bgr()[g.astype(np.int64)])


def test_area_shrink_of_constant_blocks_is_the_block_value(gpu_lib):
    say("case: imaging closed form, 2 x 3 area shrink")
    rng = np.random.default_rng(1)
    blocks = rng.integers(0, 256, (512, 512)).astype(np.float32)
    blocks[0, 0], blocks[0, 1] = 0, 255
    src = np.repeat(np.repeat(blocks, 2, axis=0), 3, axis=1)
    got = gpu_images(np.ascontiguousarray(src), 1536)
    assert np.array_equal(got[0], viridis_bgr()[blocks.astype(np.int64)])


def test_flat_slice_is_the_first_colour(gpu_lib):
    say("case: imaging closed form, flat slice")
    got = gpu_images(np.full((512, 1024), 7.0, np.float32), 512)
    assert got.shape[0] == 2
    assert np.all(got == viridis_bgr()[0])
