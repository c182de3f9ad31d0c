"""The synthetic input fields (accfft_amd/fields.py) have the spectra they
claim: checked with numpy's FFT on small grids (CPU only)."""
import numpy as np
import pytest

from accfft_amd import fields


def test_band_limited_spectrum():
    n = (32, 24, 20)
    f = fields.band_limited(n, 7, slope=-5.0 / 3.0, kmax=6.0)
    assert f.shape == n and f.dtype == np.float64
    assert abs(np.sqrt(np.mean(f * f)) - 1.0) < 1e-12
    X = np.fft.rfftn(f)
    k0 = np.fft.fftfreq(n[0], 1.0 / n[0])
    k1 = np.fft.fftfreq(n[1], 1.0 / n[1])
    k2 = np.arange(n[2] // 2 + 1)
    kabs = np.sqrt(k0[:, None, None] ** 2 + k1[None, :, None] ** 2 + k2[None, None, :] ** 2)
    inside = (kabs >= 1.0) & (kabs <= 6.0)
    # nothing outside the shell 1 <= |k| <= 6 (the mean is zero too)
    assert np.abs(X[~inside]).max() < 1e-12 * np.abs(X).max()
    # off the kz = 0 plane (where the real transform keeps the Hermitian
    # part of the random half spectrum) |X| is exactly c |k|^(-5/3)
    sel = inside.copy()
    sel[:, :, 0] = False
    ratio = np.abs(X[sel]) / kabs[sel] ** (-5.0 / 3.0)
    assert ratio.std() / ratio.mean() < 1e-10
    # random phases: not all equal
    assert np.angle(X[sel]).std() > 1.0


def test_band_limited_default_band_and_slope():
    n = (16, 16, 16)
    f = fields.band_limited(n, 3, slope=-2.0)
    X = np.fft.rfftn(f)
    k = np.fft.fftfreq(16, 1.0 / 16)
    kabs = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + np.arange(9)[None, None, :] ** 2)
    assert np.abs(X[kabs > 4.0]).max() < 1e-12 * np.abs(X).max()
    sel = (kabs >= 1) & (kabs <= 4)
    sel[:, :, 0] = False
    r = np.abs(X[sel]) * kabs[sel] ** 2
    assert r.std() / r.mean() < 1e-10


def test_sines_noise_spectrum():
    n = (16, 16, 16)
    f = fields.sines_noise(n, 1, sigma=0.0)
    X = np.fft.rfftn(f)
    a = 16 ** 3 / 2
    assert abs(X[1, 0, 0] + 1j * a) < 1e-9 and abs(X[0, 1, 0] + 1j * a) < 1e-9
    assert abs(X[0, 0, 1] + 1j * a) < 1e-9 and abs(X[15, 0, 0] - 1j * a) < 1e-9
    g = fields.sines_noise(n, 1)
    assert 0.5e-3 < np.std(g - f) < 2e-3


def test_seeded_and_distinct():
    n = (8, 8, 8)
    for make in (fields.normal, fields.dynamic_range, fields.band_limited):
        assert np.array_equal(make(n, 5), make(n, 5))
        assert not np.array_equal(make(n, 5), make(n, 6))
    d = fields.dynamic_range((64, 64, 64), 2)
    assert np.abs(d).max() / np.median(np.abs(d)) > 1e4
    c = fields.complex_of(fields.normal, n, 4)
    assert np.iscomplexobj(c) and not np.allclose(c.real, c.imag)
    s, s2 = fields.checksum(fields.normal(n, 5))
    assert s2 > 0


def test_bad_grid():
    with pytest.raises(ValueError):
        fields.normal((4, 4), 1)
