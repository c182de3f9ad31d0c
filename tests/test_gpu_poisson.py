"""GPU parity of the step4 Poisson solve (filter fused into the x pass) and
the standalone poisson_fourier_filter_gpu against the FP64 oracle
(steps/step4/kernels.cu:24-146, steps/step4/step4_gpu.cpp:520-640).
Tolerances: rel-L2 <= 1e-12 in FP64, <= 1e-5 in FP32."""
import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (12, 10, 18), (15, 9, 7), (64, 64, 64), (128, 96, 80)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_poisson_solve(dev, orc, shape, method, prec):
    import torch
    import accfft_amd as af
    print(f"poisson_solve {shape} method {method} {prec}", flush=True)
    rdt = torch.float64 if prec == "f64" else torch.float32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    fp = af.F64 if prec == "f64" else af.F32
    rho = rng(51).standard_normal(shape)
    rd = torch.tensor(rho, dtype=rdt, device=dev)
    keep = rd.clone()
    h = torch.empty((shape[0], shape[1], shape[2] // 2 + 1), dtype=cdt, device=dev)
    phi = torch.empty_like(rd)
    plan = af.Plan("r2c", shape, rd, h, fp)
    plan.poisson_solve(rd, phi, method)
    torch.cuda.synchronize()
    ref = orc.poisson_solve(rd.cpu().numpy().astype(np.float64), method)
    err = rel_l2(phi.cpu().numpy(), ref)
    print(f"  rel-L2 {err:.3e}")
    assert err < (1e-12 if prec == "f64" else 1e-5)
    assert torch.equal(rd, keep)  # rho is not modified
    plan.destroy()


@pytest.mark.parametrize("method", [0, 1])
def test_poisson_sine_512_fp32(dev, method):
    """The step4 sine case at 512^3 FP32 against its closed form."""
    import torch
    import accfft_amd as af
    n = (512, 512, 512)
    ax = torch.arange(512, device=dev, dtype=torch.float64) / 512
    s = torch.sin(2 * np.pi * ax)
    rho64 = s[:, None, None] * s[None, :, None] * s[None, None, :]
    rho = rho64.float()
    h = torch.empty((512, 512, 257), dtype=torch.complex64, device=dev)
    phi = torch.empty_like(rho)
    plan = af.Plan("r2c", n, rho, h, af.F32)
    plan.poisson_solve(rho, phi, method)
    ev = -12 * np.pi ** 2 if method == 1 else 3 * 2 * (np.cos(2 * np.pi / 512) - 1) * 512 ** 2
    exact = rho64 / ev
    err = (torch.linalg.vector_norm(phi.double() - exact) / torch.linalg.vector_norm(exact)).item()
    print(f"poisson 512^3 fp32 method {method}: rel-L2 {err:.3e}")
    assert err < 1e-5
    plan.destroy()


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_poisson_filter_block(dev, orc, method, prec):
    import torch
    import accfft_amd as af
    n = (12, 10, 18)
    h = orc.r2c(rng(52).standard_normal(n))
    blk = np.ascontiguousarray(h[3:9, 2:7, 4:10])
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    d = torch.tensor(blk, dtype=cdt, device=dev)
    before = d.cpu().numpy().astype(np.complex128)  # the filter works in place
    af.poisson_fourier_filter_gpu(d, n, blk.shape, (3, 2, 4), method)
    torch.cuda.synchronize()
    ref = orc.poisson_filter(before, n, (3, 2, 4), method)
    err = rel_l2(d.cpu().numpy(), ref)
    assert err < (1e-14 if prec == "f64" else 1e-6), err
    # the (0,0,0) mode of a block that holds it is zeroed
    d0 = torch.ones((2, 2, 2), dtype=cdt, device=dev)
    af.poisson_fourier_filter_gpu(d0, n, (2, 2, 2), (0, 0, 0), method)
    torch.cuda.synchronize()
    assert d0[0, 0, 0].item() == 0 and d0[1, 1, 1].item() != 0
