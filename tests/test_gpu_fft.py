"""GPU parity of the 3-D r2c / c2r / c2c transforms against the FP64 oracle.

Restated reference paths: src/accfft_gpu.cpp:444-623 (r2c/c2r),
:943-1136 (c2c).  Tolerances: relative L2 <= 1e-12 (FP64), <= 1e-5 (FP32)
against the FP64 oracle, as required of the drop-in paths.
"""
import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

SHAPES = [
    (8, 6, 5), (12, 10, 9), (16, 16, 16), (30, 20, 17), (7, 9, 11), (1, 1, 1), (2, 3, 4),
    (64, 64, 64), (32, 64, 128), (16, 256, 8), (128, 32, 512), (4, 4096, 2), (2048, 2, 16),
    (3, 5, 1024), (96, 40, 50),
]
TOL = {"f64": 1e-12, "f32": 1e-5}


def _t(dev, a, prec):
    import torch
    dt = torch.float64 if prec == "f64" else torch.float32
    return torch.tensor(a, dtype=dt, device=dev)


def _tc(dev, a, prec):
    import torch
    dt = torch.complex128 if prec == "f64" else torch.complex64
    return torch.tensor(a, dtype=dt, device=dev)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_c2r(dev, orc, shape, prec):
    import torch
    import accfft_amd as af
    x = rng(hash(shape) % 1000).standard_normal(shape)
    n = list(shape)
    nh = n[2] // 2 + 1
    print(f"r2c/c2r shape={shape} prec={prec}", flush=True)
    xd = _t(dev, x, prec)
    outd = torch.empty((n[0], n[1], nh), dtype=torch.complex128 if prec == "f64" else torch.complex64,
                       device=dev)
    fp = af.F64 if prec == "f64" else af.F32
    plan = af.Plan("r2c", n, xd, outd, fp)
    plan.execute_r2c(xd, outd)
    torch.cuda.synchronize()
    ref = orc.r2c(x)
    assert rel_l2(outd.cpu().numpy(), ref) < TOL[prec]
    # c2r of the oracle spectrum (non-Hermitian DC/Nyquist imag parts included)
    h = ref + 0.0
    hd = _tc(dev, h, prec)
    back = torch.empty(shape, dtype=xd.dtype, device=dev)
    plan.execute_c2r(hd, back)
    torch.cuda.synchronize()
    refb = orc.c2r(h, n)
    assert rel_l2(back.cpu().numpy(), refb) < TOL[prec]
    assert rel_l2(back.cpu().numpy() / np.prod(n), x) < TOL[prec]
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(8, 6, 5), (16, 16, 16), (30, 20, 17), (64, 32, 128), (5, 256, 7)])
def test_c2c(dev, orc, shape, prec):
    import torch
    import accfft_amd as af
    r = rng(7)
    x = r.standard_normal(shape) + 1j * r.standard_normal(shape)
    n = list(shape)
    print(f"c2c shape={shape} prec={prec}", flush=True)
    xd = _tc(dev, x, prec)
    out = torch.empty_like(xd)
    plan = af.Plan("c2c", n, xd, out, af.F64 if prec == "f64" else af.F32)
    plan.execute_c2c(af.ACCFFT_FORWARD, xd, out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.c2c(x, -1)) < TOL[prec]
    back = torch.empty_like(xd)
    src = out.clone()
    plan.execute_c2c(af.ACCFFT_BACKWARD, src, back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy() / np.prod(n), x) < TOL[prec]
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(3, 7, 32), (4, 5, 256), (3, 3, 2048), (2, 3, 4096)])
def test_z_lengths(dev, orc, shape, prec):
    """The power-of-two contiguous-line lengths the shapes above leave out
    (32, 256, 2048, 4096; odd line counts leave one unpaired real line):
    r2c, c2r, c2c and the z derivative pass against the FP64 oracle."""
    import torch
    import accfft_amd as af
    print(f"z lengths shape={shape} prec={prec}", flush=True)
    n = list(shape)
    fp = af.F64 if prec == "f64" else af.F32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    x = rng(91).standard_normal(shape)
    xd = _t(dev, x, prec)
    h = torch.empty((n[0], n[1], n[2] // 2 + 1), dtype=cdt, device=dev)
    plan = af.Plan("r2c", n, xd, h, fp, af.ACCFFT_ESTIMATE)
    plan.execute_r2c(xd, h)
    torch.cuda.synchronize()
    ref = orc.r2c(x)
    assert rel_l2(h.cpu().numpy(), ref) < TOL[prec]
    back = torch.empty_like(xd)
    plan.execute_c2r(_tc(dev, ref, prec), back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy(), orc.c2r(ref, n)) < TOL[prec]
    gz = torch.empty_like(xd)
    plan.grad(xd, None, None, gz, mask=4)
    torch.cuda.synchronize()
    assert rel_l2(gz.cpu().numpy(), orc.grad(x, 4)[2]) < TOL[prec]
    plan.destroy()
    c = x + 1j * rng(92).standard_normal(shape)
    cd = _tc(dev, c, prec)
    co = torch.empty_like(cd)
    pc = af.Plan("c2c", n, cd, co, fp, af.ACCFFT_ESTIMATE)
    pc.execute_c2c(af.ACCFFT_FORWARD, cd, co)
    torch.cuda.synchronize()
    assert rel_l2(co.cpu().numpy(), orc.c2c(c, -1)) < TOL[prec]
    pc.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(8, 6, 5), (30, 20, 17), (512, 8, 16), (6, 512, 24),
                                   (1024, 4, 12), (4, 1024, 10), (2048, 3, 8), (3, 2048, 6),
                                   (16, 32, 512), (2, 4, 2048)])
def test_c2c_inplace(dev, orc, shape, prec):
    """In-place c2c (data == data_out at plan time and at execution), the
    INPLACE build of steps/step3/step3_gpu.cpp:52-54, 61-62, 83-84, 99-100:
    forward and backward each against the FP64 oracle.  The 512-, 1024- and
    2048-point strided axes run the two-level (and, at 2048, persistent)
    tiles in place; the contiguous axes run the z pass in place."""
    import torch
    import accfft_amd as af
    r = rng(17)
    x = r.standard_normal(shape) + 1j * r.standard_normal(shape)
    n = list(shape)
    print(f"c2c in place shape={shape} prec={prec}", flush=True)
    buf = _tc(dev, x, prec)
    plan = af.Plan("c2c", n, buf, buf, af.F64 if prec == "f64" else af.F32)
    assert plan.inplace
    plan.execute_c2c(af.ACCFFT_FORWARD, buf, buf)
    torch.cuda.synchronize()
    ref = orc.c2c(x, -1)
    assert rel_l2(buf.cpu().numpy(), ref) < TOL[prec]
    # backward in place on the oracle spectrum
    buf.copy_(_tc(dev, ref, prec))
    plan.execute_c2c(af.ACCFFT_BACKWARD)  # NULL pointers: the plan's own buffer
    torch.cuda.synchronize()
    assert rel_l2(buf.cpu().numpy(), orc.c2c(ref, +1)) < TOL[prec]
    assert rel_l2(buf.cpu().numpy() / np.prod(n), x) < TOL[prec]
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(3, 128, 256), (2, 512, 512), (5, 256, 1024), (2, 2048, 2048),
                                   (7, 1024, 512), (1, 2048, 256)])
def test_fused_z_y_first_level(dev, orc, shape, prec):
    """FP32 plans whose z pass carries the first radix-8 level of the y
    transform (N1 = 8 M, M = 16 ... 256, z lengths 256 ... 2048; FP64 plans
    keep separate passes, which are faster there): r2c,
    c2r (also with PRESERVE_INPUT), laplace and the slow grad against the
    oracle, and against the same plan with the fusion switched off
    (tuning choice -3 on the y axis)."""
    import torch
    import accfft_amd as af
    n = list(shape)
    nh = n[2] // 2 + 1
    print(f"fused z+y shape={shape} prec={prec}", flush=True)
    fp = af.F64 if prec == "f64" else af.F32
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    x = rng(31).standard_normal(shape)
    xd = _t(dev, x, prec)
    h = torch.empty((n[0], n[1], nh), dtype=cdt, device=dev)
    plan = af.Plan("r2c", n, xd, h, fp, af.ACCFFT_ESTIMATE | af.PRESERVE_INPUT)
    fused = prec == "f32"
    assert plan.tuning()[1] != -3
    plan.execute_r2c(xd, h)
    torch.cuda.synchronize()
    ref = orc.r2c(x)
    assert rel_l2(h.cpu().numpy(), ref) < TOL[prec]
    hd = _tc(dev, ref, prec)
    back = torch.empty_like(xd)
    plan.execute_c2r(hd, back)  # PRESERVE_INPUT: hd is kept
    torch.cuda.synchronize()
    refb = orc.c2r(ref, n)
    assert rel_l2(back.cpu().numpy(), refb) < TOL[prec]
    assert rel_l2(hd.cpu().numpy(), ref) < TOL[prec]
    lap = torch.empty_like(xd)
    plan.laplace(xd, lap)
    gy = torch.empty_like(xd)
    plan.grad(xd, None, gy, None, mask=2, slow=True)
    torch.cuda.synchronize()
    assert rel_l2(lap.cpu().numpy(), orc.laplace(x)) < TOL[prec]
    assert rel_l2(gy.cpu().numpy(), orc.grad(x, 2, slow=True)[1]) < TOL[prec]
    if not fused:  # no fused z pass at this length: nothing to switch off
        with pytest.raises(af.AccfftError):
            plan.set_tuning(1, -3)
        plan.destroy()
        return
    # the separate passes give the same transform
    plan.set_tuning(1, -3)
    assert plan.tuning()[1] == -3
    h2 = torch.empty_like(h)
    plan.execute_r2c(xd, h2)
    torch.cuda.synchronize()
    assert rel_l2(h2.cpu().numpy(), ref) < TOL[prec]
    plan.set_tuning(1, -1)
    assert plan.tuning()[1] != -3
    plan.destroy()


# Powers of three: radix-9 stages of the run-time schedule (729 = 9 9 9,
# 243 = 9 9 3) in the LDS tiles of every axis, and in the global radix passes
# of lines longer than one tile (19683 = 9^4 3 in both precisions, 6561 =
# 9^4 in FP64).  The oracle factors its lines with radix 3 only.
P3_SHAPES = [(27, 81, 243), (9, 729, 18), (729, 3, 9), (4, 3, 2187), (2, 3, 19683), (6561, 2, 4)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", P3_SHAPES)
def test_powers_of_three(dev, orc, shape, prec):
    import torch
    import accfft_amd as af
    print(f"powers of three shape={shape} prec={prec}", flush=True)
    n = list(shape)
    fp = af.F64 if prec == "f64" else af.F32
    x = rng(sum(shape)).standard_normal(shape)
    xd = _t(dev, x, prec)
    h = torch.empty((n[0], n[1], n[2] // 2 + 1),
                    dtype=torch.complex128 if prec == "f64" else torch.complex64, device=dev)
    plan = af.Plan("r2c", n, xd, h, fp)
    plan.execute_r2c(xd, h)
    torch.cuda.synchronize()
    ref = orc.r2c(x)
    assert rel_l2(h.cpu().numpy(), ref) < TOL[prec]
    back = torch.empty_like(xd)
    plan.execute_c2r(h, back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy() / np.prod(n), x) < TOL[prec]
    plan.destroy()
    r = rng(sum(shape) + 1)
    z = r.standard_normal(shape) + 1j * r.standard_normal(shape)
    zd = _tc(dev, z, prec)
    out = torch.empty_like(zd)
    pc = af.Plan("c2c", n, zd, out, fp)
    pc.execute_c2c(af.ACCFFT_BACKWARD, zd, out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.c2c(z, +1)) < TOL[prec]
    pc.destroy()
