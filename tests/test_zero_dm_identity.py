"""The u8 clean kernel's integer zero-DM: for a row mean m (float32) and a
byte x, the float expression (uint8)trunc(max((float)x - m, 0)) of the
reference (src/rfi/rfimarker.cu:172-305) equals max(x - ceil(m), 0) whenever
m is an integer or frac(m) >= 2^-16; the kernel takes the float path for the
other means.  Checked here for every byte value against float32 arithmetic."""
import numpy as np


def float_path(x, m):
    f = x.astype(np.float32) - np.float32(m)
    f = np.where(f < 0, np.float32(0), f)
    return f.astype(np.uint32)


def int_path(x, m):
    m = np.float32(m)
    fl = np.floor(m)
    k = int(fl) + (1 if m - fl > 0 else 0)
    return np.maximum(x.astype(np.int64) - k, 0).astype(np.uint32)


def fast_ok(m):
    m = np.float32(m)
    fr = m - np.floor(m)
    return fr == 0 or fr >= np.float32(2.0 ** -16)


def test_identity_on_the_guarded_means():
    x = np.arange(256, dtype=np.uint32)
    rng = np.random.default_rng(0)
    ms = [np.float32(v) for v in range(256)]
    for k in range(256):
        base = np.float32(k)
        for fr in (2.0 ** -16, 2.0 ** -16 + 2.0 ** -22, 0.25, 0.5, 0.999, 1 - 2.0 ** -15):
            ms.append(np.float32(base + np.float32(fr)))
        ms.extend(np.float32(k + f) for f in rng.random(8))
    # the means the kernel can see: s / n for n <= 4096 channels
    for n in (1, 3, 7, 255, 4095, 4096):
        s = rng.integers(0, 255 * n + 1, 300)
        ms.extend(np.float32(np.float32(v) / np.float32(n)) for v in s)
    checked = 0
    for m in ms:
        if not (0 <= m < 256) or not fast_ok(m):
            continue
        assert np.array_equal(float_path(x, m), int_path(x, m)), float(m)
        checked += 1
    assert checked > 3000


def test_the_guard_is_needed():
    # a mean just above an integer, within rounding of x - m: the float path
    # rounds up to the next integer for large x, the byte subtract does not
    x = np.arange(256, dtype=np.uint32)
    m = np.float32(3.0) + np.float32(2.0 ** -20)
    assert not fast_ok(m)
    assert not np.array_equal(float_path(x, m), int_path(x, m))
