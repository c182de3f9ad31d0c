"""Seeded random shapes (every axis 1 ... 70, primes and odd lengths
included) through r2c, c2r and c2c in both precisions, and random pencil
grids over them, against the oracle: the edge cases a fixed list of shapes
can miss (axes of one point, Nyquist bins of odd lengths, ragged pencil
blocks)."""
import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

_R = np.random.default_rng(20261203)
SHAPES = [tuple(int(v) for v in _R.integers(1, 71, size=3)) for _ in range(24)]
SHAPES += [(1, 1, 1), (1, 5, 1), (7, 1, 2), (2, 3, 1), (1, 1, 9)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_random_shape_vs_oracle(dev, orc, shape, prec):
    import torch
    import accfft_amd as af
    print(f"random shape {shape} {prec}", flush=True)
    fp, rdt, cdt, tol = ((af.F64, torch.float64, torch.complex128, 1e-12) if prec == "f64"
                         else (af.F32, torch.float32, torch.complex64, 1e-5))
    n = list(shape)
    x = rng(sum(shape)).standard_normal(shape)
    xd = torch.tensor(x, dtype=rdt, device=dev)
    h = torch.empty((n[0], n[1], n[2] // 2 + 1), dtype=cdt, device=dev)
    p = af.Plan("r2c", n, xd, h, fp, af.ACCFFT_ESTIMATE)
    p.execute_r2c(xd, h)
    torch.cuda.synchronize()
    ref = orc.r2c(x)
    assert rel_l2(h.cpu().numpy(), ref) < tol
    back = torch.empty_like(xd)
    p.execute_c2r(torch.tensor(ref, dtype=cdt, device=dev), back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy(), orc.c2r(ref, n)) < tol
    p.destroy()
    c = x + 1j * rng(sum(shape) + 1).standard_normal(shape)
    cd = torch.tensor(c, dtype=cdt, device=dev)
    co = torch.empty_like(cd)
    q = af.Plan("c2c", n, cd, co, fp, af.ACCFFT_ESTIMATE)
    q.execute_c2c(af.ACCFFT_FORWARD, cd, co)
    torch.cuda.synchronize()
    assert rel_l2(co.cpu().numpy(), orc.c2c(c, -1)) < tol
    q.execute_c2c(af.ACCFFT_BACKWARD, co, cd)
    torch.cuda.synchronize()
    assert rel_l2(cd.cpu().numpy() / np.prod(n), c) < tol
    q.destroy()
