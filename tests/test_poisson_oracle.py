"""Oracle pins for the step4 Poisson Fourier filter
(steps/step4/kernels.cu:24-110, steps/step4/step4_gpu.cpp:520-640).

The reference holds no stored vectors for it; its step4 program checks the
sine eigenfunction case.  These tests pin both methods with what the
mathematics fixes, independently of the oracle's symbol code:
  * the sine case in closed form (method 1: the continuous eigenvalue
    -12 pi^2; method 0: the 7-point discrete one),
  * method 0: the periodic 7-point Laplacian (spacing 1/N_d) of the solution
    equals rho minus its mean, on even, odd and non-cubic grids,
  * method 1: numpy's FFT with -4 pi^2 |k|^2 applied to the solution gives
    rho minus its mean (Nyquist modes included: (+N/2)^2 = (-N/2)^2),
  * the block form (offsets ostart) equals the slice of the full filter.
"""
import numpy as np
import pytest

from conftest import rel_l2, rng


def _grid(n):
    return np.meshgrid(*[np.arange(m) / m for m in n], indexing="ij")


@pytest.mark.parametrize("n", [(16, 16, 16), (12, 10, 18), (15, 9, 7)])
def test_sine_closed_form(orc, n):
    x, y, z = _grid(n)
    rho = np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    phi1 = orc.poisson_solve(rho, method=1)
    # step4_gpu.cpp:111-115: the exact solution
    assert rel_l2(phi1, -rho / (3 * 4 * np.pi ** 2)) < 1e-12
    phi0 = orc.poisson_solve(rho, method=0)
    ev = sum(2 * (np.cos(2 * np.pi / m) - 1) * m * m for m in n)
    assert rel_l2(phi0, rho / ev) < 1e-12


def _lap7(phi):
    n = phi.shape
    out = np.zeros_like(phi)
    for d in range(3):
        out += (np.roll(phi, 1, d) - 2 * phi + np.roll(phi, -1, d)) * n[d] ** 2
    return out


@pytest.mark.parametrize("n", [(16, 16, 16), (12, 10, 18), (15, 9, 7), (8, 6, 5)])
def test_discrete_method_inverts_7point_laplacian(orc, n):
    rho = rng(41).standard_normal(n)
    phi = orc.poisson_solve(rho, method=0)
    assert abs(phi.mean()) < 1e-13
    assert rel_l2(_lap7(phi), rho - rho.mean()) < 1e-11


@pytest.mark.parametrize("n", [(16, 16, 16), (12, 10, 18), (15, 9, 7), (8, 6, 5)])
def test_continuous_method_inverts_spectral_laplacian(orc, n):
    rho = rng(42).standard_normal(n)
    phi = orc.poisson_solve(rho, method=1)
    k2 = sum(np.meshgrid(*[(np.fft.fftfreq(m) * m) ** 2 for m in n], indexing="ij"))
    lap = np.fft.ifftn(-4 * np.pi ** 2 * k2 * np.fft.fftn(phi)).real
    assert abs(phi.mean()) < 1e-13
    assert rel_l2(lap, rho - rho.mean()) < 1e-11


@pytest.mark.parametrize("method", [0, 1])
def test_filter_block_offsets(orc, method):
    n = (12, 10, 18)
    h = orc.r2c(rng(43).standard_normal(n))
    full = orc.poisson_filter(h, n, (0, 0, 0), method)
    blk = h[3:9, 2:7, 4:10]
    part = orc.poisson_filter(blk, n, (3, 2, 4), method)
    assert np.array_equal(part, full[3:9, 2:7, 4:10])
    assert full[0, 0, 0] == 0
