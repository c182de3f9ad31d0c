"""The dynamic-engine stages form run-time quotients (butterfly index to
line and index, index modulo the span) as q = trunc(float(b) * (1.0f / d))
with one correction step (accfft_amd/csrc/fft_device.hpp: fdivmod) instead
of integer division.  Exhaustive check in IEEE float32 (numpy rounds the
reciprocal, the product and the conversion as the device does): for every
divisor 1 <= d <= 10240 and every dividend 0 <= b < 10240 (a dynamic tile
holds at most 160 KiB / 2 buffers / 8 bytes = 10240 elements, so every
butterfly and element index is below that), the corrected quotient and
remainder equal b // d and b % d."""
import numpy as np


def _fdivmod(b, d):
    rcp = np.float32(1.0) / np.float32(d)
    q = (b.astype(np.float32) * rcp).astype(np.int64)  # truncation toward zero
    r = b - q * d
    lo = r < 0
    hi = r >= d
    q = np.where(lo, q - 1, np.where(hi, q + 1, q))
    r = np.where(lo, r + d, np.where(hi, r - d, r))
    return q, r


def test_fdivmod_exhaustive_tile_range():
    b = np.arange(10240, dtype=np.int64)
    for d in range(1, 10241):
        q, r = _fdivmod(b, d)
        assert np.array_equal(q, b // d), d
        assert np.array_equal(r, b % d), d


def test_fdivmod_needs_the_correction():
    # without the correction step the float quotient is wrong for some
    # pairs in range, so the check above is not vacuous
    b = np.arange(10240, dtype=np.int64)
    wrong = 0
    for d in range(1, 10241):
        rcp = np.float32(1.0) / np.float32(d)
        q = (b.astype(np.float32) * rcp).astype(np.int64)
        wrong += int(np.count_nonzero(q != b // d))
    assert wrong > 0
