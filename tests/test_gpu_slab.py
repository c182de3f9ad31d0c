"""Slab-share plans: one rank's r2c / c2r on a P x 1 grid with the y pass
storing the packed exchange blocks and the x pass reading the received ones.

Against the oracle's composition (z/y r2c, local_transpose_col pack, x c2c)
at small sizes in FP64; at the full 256 x 2048 x 2048 FP32 share through
the round trip and a field whose forward output has a closed form.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

CASES = [((8, 16, 10), 4), ((16, 32, 24), 4), ((64, 64, 64), 8), ((24, 16, 33), 2),
         ((12, 10, 6), 1), ((8, 4, 6), 4)]  # the last: one y row per rank


def _plan(N, P, fp, a, h):
    import accfft_amd as af
    return af.Plan("r2c_slab", N, a, h, fp, af.ACCFFT_MEASURE, nprocs=P)


@pytest.mark.parametrize("N,P", CASES)
def test_slab_vs_oracle_fp64(dev, orc, N, P):
    import torch
    import accfft_amd as af
    print(f"slab FP64 {N} P={P}", flush=True)
    n0, n1, Nh = N[0] // P, N[1] // P, N[2] // 2 + 1
    a = rng(70).standard_normal((n0, N[1], N[2]))
    ad = torch.tensor(a, device=dev)
    h = torch.empty((N[0], n1, Nh), dtype=torch.complex128, device=dev)
    plan = _plan(N, P, af.F64, ad, h)
    assert plan.isize == (n0, N[1], N[2]) and plan.osize == (N[0], n1, Nh)
    plan.execute_r2c(ad, h)
    torch.cuda.synchronize()
    ref = orc.slab_r2c(a, P)
    assert rel_l2(h.cpu().numpy(), ref) < 1e-12
    back = torch.empty_like(ad)
    hc = h.clone()
    plan.execute_c2r(hc, back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy(), orc.slab_c2r(h.cpu().numpy(), N[2], P)) < 1e-12
    assert rel_l2(back.cpu().numpy(), np.prod(N) * a) < 1e-12
    plan.destroy()


def _hip():
    """The HIP runtime this process already loaded (torch's)."""
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                return ctypes.CDLL(line.split()[-1])
    raise RuntimeError("libamdhip64 not loaded")


def test_slab_exchange_hook(dev):
    """A hook that copies the blocks (the one-rank all-to-all) gives the same
    bits as the identity exchange, and is called once per direction."""
    import torch
    import accfft_amd as af
    N, P = (32, 64, 40), 4
    n0, n1, Nh = N[0] // P, N[1] // P, N[2] // 2 + 1
    ad = torch.randn((n0, N[1], N[2]), dtype=torch.float64, device=dev)
    h0 = torch.empty((N[0], n1, Nh), dtype=torch.complex128, device=dev)
    h1 = torch.empty_like(h0)
    plan = _plan(N, P, af.F64, ad, h0)
    plan.execute_r2c(ad, h0)
    hip = _hip()
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int, ctypes.c_void_p]
    calls = []

    def hook(send, recv, nbytes, nprocs, direction, stream):
        calls.append((nbytes, nprocs, direction))
        return 0 if hip.hipMemcpyAsync(recv, send, nbytes * nprocs, 3, stream) == 0 else 5

    plan.set_exchange(hook)
    timer = [0.0] * 5
    plan.execute_r2c(ad, h1, timer=timer)
    torch.cuda.synchronize()
    assert torch.equal(h0, h1)
    # doc/timing.md: the exchange in timer[2], the local FFT in timer[4]
    assert timer[2] > 0 and timer[4] > 0 and timer[0] == timer[1] == timer[3] == 0, timer
    b0, b1 = torch.empty_like(ad), torch.empty_like(ad)
    hc = h1.clone()
    plan.execute_c2r(h1, b1)
    plan.set_exchange(None)
    timer = [0.0] * 5
    plan.execute_c2r(hc, b0, timer=timer)
    torch.cuda.synchronize()
    assert torch.equal(b0, b1)
    assert timer[2] == 0 and timer[4] > 0, timer  # identity exchange: no communication
    blk = n0 * n1 * Nh * 16
    assert calls == [(blk, P, -1), (blk, P, 1)]
    plan.destroy()


def test_slab_full_share_fp32(dev):
    """256 x 2048 x 2048 FP32 share of 2048^3 on 8 ranks: round trip, and the
    forward of cos(2 pi (ky y + kz z) / 2048) against its closed form (only
    block q = ky // 256 is non-zero before the x transform, constant over
    its 256 x rows, so the output is one column of box-function spectra)."""
    import torch
    import accfft_amd as af
    N, P = (2048, 2048, 2048), 8
    n0, n1, Nh = 256, 256, 1025
    print("slab FP32 256x2048x2048 P=8", flush=True)
    ky, kz = 700, 3
    y = torch.arange(N[1], device=dev, dtype=torch.float64).view(1, -1, 1)
    z = torch.arange(N[2], device=dev, dtype=torch.float64).view(1, 1, -1)
    a = torch.cos(2 * np.pi * (ky * y + kz * z) / N[1]).float().expand(n0, -1, -1).contiguous()
    h = torch.empty((N[0], n1, Nh), dtype=torch.complex64, device=dev)
    plan = _plan(N, P, af.F32, a, h)
    plan.execute_r2c(a, h)
    torch.cuda.synchronize()
    q, yl = ky // n1, ky % n1
    k = torch.arange(N[0], device=dev, dtype=torch.float64)
    X = torch.arange(q * n0, (q + 1) * n0, device=dev, dtype=torch.float64)
    box = torch.exp(-2j * np.pi * k.view(-1, 1) * X.view(1, -1) / N[0]).sum(dim=1)
    col = box * (N[1] * N[2] / 2)
    got = h[:, yl, kz].to(torch.complex128)
    err_col = ((got - col).abs().norm() / col.abs().norm()).item()
    h[:, yl, kz] = 0
    rest = (h.abs().double().norm() / col.abs().norm()).item()
    print(f"  closed form: column rel-L2 {err_col:.2e}, rest {rest:.2e}")
    assert err_col < 1e-5 and rest < 1e-5
    del a, h
    torch.cuda.empty_cache()
    g = torch.Generator(device=dev).manual_seed(71)
    u = torch.randn((n0, N[1], N[2]), generator=g, device=dev)
    h = torch.empty((N[0], n1, Nh), dtype=torch.complex64, device=dev)
    plan.execute_r2c(u, h)
    back = torch.empty_like(u)
    plan.execute_c2r(h, back)
    torch.cuda.synchronize()
    back /= float(np.prod(N))
    err = ((back - u).double().norm() / u.double().norm()).item()
    print(f"  round trip rel-L2 {err:.2e}")
    assert err < 1e-5
    plan.destroy()


@pytest.mark.parametrize("N,P", [((16, 2048, 2048), 8), ((2, 2048, 2048), 1), ((64, 2048, 2048), 64)])
def test_slab_fused_first_y_level_fp32(dev, orc, N, P):
    """2048-point y and z axes in FP32: the z pass carries the first radix-8
    level of the y transform and the y pass is eight 256-point passes whose
    outputs are the rows k1 + 8 k2 of the exchange blocks (16 local rows per
    block at P = 64, one block at P = 1).  Forward and inverse against the
    oracle, and the round trip."""
    import torch
    import accfft_amd as af
    print(f"slab fused FP32 {N} P={P}", flush=True)
    n0, n1, Nh = N[0] // P, N[1] // P, N[2] // 2 + 1
    a = rng(71).standard_normal((n0, N[1], N[2]))
    ad = torch.tensor(a, dtype=torch.float32, device=dev)
    h = torch.empty((N[0], n1, Nh), dtype=torch.complex64, device=dev)
    plan = af.Plan("r2c_slab", N, ad, h, af.F32, af.ACCFFT_ESTIMATE, nprocs=P)
    plan.execute_r2c(ad, h)
    torch.cuda.synchronize()
    ref = orc.slab_r2c(a, P)
    assert rel_l2(h.cpu().numpy(), ref) < 1e-5
    back = torch.empty_like(ad)
    hc = h.clone()
    plan.execute_c2r(hc, back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy(), orc.slab_c2r(h.cpu().numpy().astype(np.complex128), N[2], P)) < 1e-5
    assert rel_l2(back.cpu().numpy() / np.prod(N), a) < 1e-5
    plan.destroy()


def test_slab_fused_exchange_hook_fp32(dev):
    """The fused 2048-point slab path through an exchange hook that copies
    the blocks: same bits as the identity exchange, forward and inverse."""
    import torch
    import accfft_amd as af
    N, P = (16, 2048, 2048), 8
    n0, n1, Nh = N[0] // P, N[1] // P, N[2] // 2 + 1
    print(f"slab fused FP32 hook {N} P={P}", flush=True)
    ad = torch.randn((n0, N[1], N[2]), dtype=torch.float32, device=dev)
    h0 = torch.empty((N[0], n1, Nh), dtype=torch.complex64, device=dev)
    h1 = torch.empty_like(h0)
    plan = af.Plan("r2c_slab", N, ad, h0, af.F32, af.ACCFFT_ESTIMATE, nprocs=P)
    plan.execute_r2c(ad, h0)
    hip = _hip()
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int, ctypes.c_void_p]

    def hook(send, recv, nbytes, nprocs, direction, stream):
        return 0 if hip.hipMemcpyAsync(recv, send, nbytes * nprocs, 3, stream) == 0 else 5

    plan.set_exchange(hook)
    plan.execute_r2c(ad, h1)
    torch.cuda.synchronize()
    assert torch.equal(h0, h1)
    b0, b1 = torch.empty_like(ad), torch.empty_like(ad)
    hc = h1.clone()
    plan.execute_c2r(h1, b1)
    plan.set_exchange(None)
    plan.execute_c2r(hc, b0)
    torch.cuda.synchronize()
    assert torch.equal(b0, b1)
    plan.destroy()


def test_slab_blocked_2048_y_persistent_fp32(dev, orc):
    """2048-point y axis with z = 64 (no fused first y level): the y pass
    runs the persistent two-level tiles over the exchange blocks, with more
    tiles than workgroups (64 x 33 columns in 8-column tiles), so some go
    through the plan's tile queue.  Against the oracle's composition
    (rel-L2 <= 1e-5 FP32), forward and inverse, twice in a row."""
    import torch
    import accfft_amd as af
    N, P = (256, 2048, 64), 4
    n0, n1, Nh = N[0] // P, N[1] // P, N[2] // 2 + 1
    print(f"slab FP32 blocked 2048-point y {N} P={P}", flush=True)
    a = rng(72).standard_normal((n0, N[1], N[2]))
    ad = torch.tensor(a, dtype=torch.float32, device=dev)
    h = torch.empty((N[0], n1, Nh), dtype=torch.complex64, device=dev)
    # the built-in rule (the tuner may pick another configuration)
    plan = af.Plan("r2c_slab", N, ad, h, af.F32, af.ACCFFT_ESTIMATE, nprocs=P)
    ref = orc.slab_r2c(a, P)
    for rep in range(2):
        plan.execute_r2c(ad, h)
        torch.cuda.synchronize()
        assert rel_l2(h.cpu().numpy(), ref) < 1e-5, rep
        back = torch.empty_like(ad)
        hc = h.clone()
        plan.execute_c2r(hc, back)
        torch.cuda.synchronize()
        assert rel_l2(back.cpu().numpy(), np.prod(N) * a) < 1e-5, rep
    plan.destroy()
