"""Axes longer than one LDS tile (FP32 > 10240 points, FP64 > 5120 points
that are not powers of two up to 4096), which cuFFT plans accept
(src/accfft_gpu.cpp:173-253, any n) and which run here as one global radix
pass per stage of the mixed-radix schedule, or, for lengths with a prime
factor above 61, as Bluestein's chirp convolution over power-of-two FFTs:
r2c, c2r and c2c with the long axis in each position, prime and
large-prime-factor lengths, and the spectral operators, against the FP64
oracle."""
import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


def _dt(prec, cplx=False):
    import torch
    if cplx:
        return torch.complex128 if prec == "f64" else torch.complex64
    return torch.float64 if prec == "f64" else torch.float32


CASES = [
    ((2, 3, 8192), "f64"), ((8192, 3, 2), "f64"), ((3, 8192, 2), "f64"),
    ((3, 2, 6000), "f64"), ((2, 5153, 3), "f64"), ((5153, 2, 2), "f64"),
    ((2, 3, 16384), "f32"), ((16384, 2, 3), "f32"), ((2, 16384, 2), "f32"),
    ((3, 2, 12000), "f32"), ((2, 3, 10243), "f32"),
    # Bluestein: primes and a large prime factor (8128 = 64 * 127)
    ((2, 2, 10007), "f64"), ((8128, 2, 3), "f64"), ((3, 12289, 2), "f32"), ((20011, 2, 2), "f32"),
]


@pytest.mark.parametrize("shape,prec", CASES)
def test_long_r2c_c2r_c2c(dev, orc, shape, prec):
    import torch
    import accfft_amd as af
    print(f"long lines {shape} {prec}", flush=True)
    n = list(shape)
    fp = af.F64 if prec == "f64" else af.F32
    x = rng(41).standard_normal(shape)
    xd = torch.tensor(x, dtype=_dt(prec), device=dev)
    h = torch.empty((n[0], n[1], n[2] // 2 + 1), dtype=_dt(prec, True), device=dev)
    plan = af.Plan("r2c", n, xd, h, fp)
    plan.execute_r2c(xd, h)
    torch.cuda.synchronize()
    ref = orc.r2c(x)
    assert rel_l2(h.cpu().numpy(), ref) < TOL[prec]
    back = torch.empty_like(xd)
    plan.execute_c2r(torch.tensor(ref, dtype=_dt(prec, True), device=dev), back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy(), orc.c2r(ref, n)) < TOL[prec]
    plan.destroy()
    c = x + 1j * rng(42).standard_normal(shape)
    cd = torch.tensor(c, dtype=_dt(prec, True), device=dev)
    co = torch.empty_like(cd)
    pc = af.Plan("c2c", n, cd, co, fp)
    pc.execute_c2c(af.ACCFFT_FORWARD, cd, co)
    torch.cuda.synchronize()
    refc = orc.c2c(c, -1)
    assert rel_l2(co.cpu().numpy(), refc) < TOL[prec]
    # in place, backward
    pc.destroy()
    buf = torch.tensor(refc, dtype=_dt(prec, True), device=dev)
    pi = af.Plan("c2c", n, buf, buf, fp)
    pi.execute_c2c(af.ACCFFT_BACKWARD, buf, buf)
    torch.cuda.synchronize()
    assert rel_l2(buf.cpu().numpy() / np.prod(n), c) < TOL[prec]
    pi.destroy()


@pytest.mark.parametrize("shape", [(4, 3, 8192), (8192, 4, 3), (3, 6000, 4), (4, 5158, 2)])
def test_long_operators_fp64(dev, orc, shape):
    """laplace, inverse laplace, grad (the 3-D form on long axes) and
    divergence against the oracle's operators."""
    import torch
    import accfft_amd as af
    print(f"long-line operators {shape}", flush=True)
    x = rng(43).standard_normal(shape)
    xd = torch.tensor(x, device=dev)
    plan = af.Plan("r2c", list(shape), xd, None, af.F64)
    out = torch.empty_like(xd)
    plan.laplace(xd, out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.laplace(x)) < 1e-12
    plan.inv_laplace(xd, out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.inv_laplace(x)) < 1e-12
    g = [torch.empty_like(xd) for _ in range(3)]
    plan.grad(xd, *g)
    torch.cuda.synchronize()
    ref = orc.grad(x, 7)
    for d in range(3):
        assert rel_l2(g[d].cpu().numpy(), ref[d]) < 1e-12, d
    plan.divergence(g[0], g[1], g[2], out)
    torch.cuda.synchronize()
    refd = orc.divergence(*[gi.cpu().numpy() for gi in g])
    assert rel_l2(out.cpu().numpy(), refd) < 1e-12
    plan.destroy()


def test_long_lines_rejected_where_unsupported(dev):
    """Slab shares need every axis in one LDS tile; the per-axis r2c of a
    strided long axis has no long-line form: both report unsupported."""
    import torch
    import accfft_amd as af
    with pytest.raises(af.AccfftError, match="unsupported"):
        af.Plan("r2c_slab", (8, 8, 8192), None, None, af.F64, nprocs=2)
    x = torch.zeros((8192, 2, 2), dtype=torch.float64, device=dev)
    o = torch.empty((4097, 2, 2), dtype=torch.complex128, device=dev)
    plan = af.Plan("r2c", (8192, 2, 2), x, None, af.F64)
    with pytest.raises(af.AccfftError, match="unsupported"):
        plan.execute_r2c_axis(0, x, o)
    plan.destroy()
