"""GPU parity of the spectral operators, per-axis transforms, axis masks and
the in-place r2c layout against the FP64 oracle.

Reference paths: src/operators_gpu.txx:72-609 (grad fast/slow, laplace,
divergence fast/slow, biharmonic), src/operators.txx:276-392 and :1270-1321
(inverse laplace / biharmonic, CPU-only upstream), src/accfft_gpu.cpp:
1212-1604 (per-axis), :471-603 (xyz masks), :99-119 (in-place plans).
"""
import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}
SHAPES = [(16, 12, 20), (7, 9, 11), (30, 20, 17), (64, 64, 64), (32, 128, 16)]


def _dt(prec):
    import torch
    return (torch.float64, torch.complex128) if prec == "f64" else (torch.float32, torch.complex64)


def _plan(n, prec, x=None, out=None, flags=2):
    import accfft_amd as af
    return af.Plan("r2c", n, x, out, af.F64 if prec == "f64" else af.F32, flags)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES)
def test_grad_div(dev, orc, shape, prec):
    import torch
    rdt, _ = _dt(prec)
    print(f"grad/div shape={shape} prec={prec}", flush=True)
    a = rng(1).standard_normal(shape)
    ad = torch.tensor(a, dtype=rdt, device=dev)
    plan = _plan(shape, prec)
    outs = [torch.empty_like(ad) for _ in range(3)]
    plan.grad(ad, *outs)
    slow = [torch.empty_like(ad) for _ in range(3)]
    plan.grad(ad, *slow, slow=True)
    torch.cuda.synchronize()
    ref = orc.grad(a)
    refs = orc.grad(a, slow=True)
    for d in range(3):
        assert rel_l2(outs[d].cpu().numpy(), ref[d]) < TOL[prec], d
        assert rel_l2(slow[d].cpu().numpy(), refs[d]) < TOL[prec], d
    # divergence of three independent fields
    f = [rng(10 + d).standard_normal(shape) for d in range(3)]
    fd = [torch.tensor(v, dtype=rdt, device=dev) for v in f]
    div = torch.empty_like(ad)
    plan.divergence(*fd, div)
    divs = torch.empty_like(ad)
    plan.divergence(*fd, divs, slow=True)
    torch.cuda.synchronize()
    assert rel_l2(div.cpu().numpy(), orc.divergence(*f)) < TOL[prec]
    assert rel_l2(divs.cpu().numpy(), orc.divergence(*f, slow=True)) < TOL[prec]
    # mask: only y
    only_y = torch.zeros_like(ad)
    plan.grad(ad, None, only_y, None, mask=0b010)
    torch.cuda.synchronize()
    assert rel_l2(only_y.cpu().numpy(), ref[1]) < TOL[prec]
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("op", ["laplace", "biharmonic", "inv_laplace", "inv_biharmonic"])
def test_spectral_ops(dev, orc, shape, prec, op):
    import torch
    rdt, _ = _dt(prec)
    print(f"{op} shape={shape} prec={prec}", flush=True)
    a = rng(2).standard_normal(shape)
    ad = torch.tensor(a, dtype=rdt, device=dev)
    out = torch.empty_like(ad)
    plan = _plan(shape, prec)
    getattr(plan, op)(ad, out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), getattr(orc, op)(a)) < TOL[prec]
    # the input is left intact
    assert np.array_equal(ad.cpu().numpy(), a.astype(ad.cpu().numpy().dtype))
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(8, 6, 5), (7, 9, 11), (16, 32, 64), (64, 16, 12)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_axis_transforms(dev, orc, shape, prec, axis):
    import torch
    rdt, cdt = _dt(prec)
    print(f"axis {axis} shape={shape} prec={prec}", flush=True)
    a = rng(3).standard_normal(shape)
    ad = torch.tensor(a, dtype=rdt, device=dev)
    hs = list(shape)
    hs[axis] = shape[axis] // 2 + 1
    h = torch.empty(hs, dtype=cdt, device=dev)
    plan = _plan(shape, prec)
    plan.execute_r2c_axis(axis, ad, h)
    torch.cuda.synchronize()
    href = orc.r2c_axis(a, axis)
    assert rel_l2(h.cpu().numpy(), href) < TOL[prec]
    back = torch.empty_like(ad)
    plan.execute_c2r_axis(axis, h, back)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy(), orc.c2r_axis(href, shape, axis)) < TOL[prec]
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("xyz", [0b101, 0b110, 0b100, 0b011, 0b001])
def test_xyz_masks(dev, orc, prec, xyz):
    import torch
    rdt, cdt = _dt(prec)
    shape = (8, 12, 10)
    nh = 6
    a = rng(4).standard_normal(shape)
    if xyz & 4:
        ad = torch.tensor(a, dtype=rdt, device=dev)
        h = torch.empty((8, 12, nh), dtype=cdt, device=dev)
        # out-of-place plan: data != data_out (src/accfft_gpu.cpp:99-103)
        plan = _plan(shape, prec, ad, h)
        plan.execute_r2c(ad, h, xyz=xyz)
        torch.cuda.synchronize()
        ref = orc.r2c(a, xyz)
        assert rel_l2(h.cpu().numpy(), ref) < TOL[prec]
        back = torch.empty_like(ad)
        plan.execute_c2r(h, back, xyz=xyz)
        torch.cuda.synchronize()
        assert rel_l2(back.cpu().numpy(), orc.c2r(ref, shape, xyz)) < TOL[prec]
    else:
        # z skipped: the transform runs in place on the input buffer,
        # reinterpreted as complex [N0][N1][Nh] (src/accfft_gpu.cpp:488-489)
        c = rng(5).standard_normal((8, 12, nh)) + 1j * rng(6).standard_normal((8, 12, nh))
        cd = torch.tensor(c, dtype=cdt, device=dev)
        plan = _plan(shape, prec, cd, cd)
        plan.execute_r2c(cd, None, xyz=xyz)
        torch.cuda.synchronize()
        ref = c.copy()
        if xyz & 2:
            ref = np.fft.fft(ref, axis=1)
        if xyz & 1:
            ref = np.fft.fft(ref, axis=0)
        assert rel_l2(cd.cpu().numpy(), ref) < TOL[prec]
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(8, 6, 10), (16, 16, 16), (5, 7, 9), (64, 32, 128)])
def test_inplace_r2c(dev, orc, shape, prec):
    """In-place plan: rows padded to 2*(N2/2+1) reals (steps/step2)."""
    import torch
    rdt, cdt = _dt(prec)
    n0, n1, n2 = shape
    nh = n2 // 2 + 1
    a = rng(7).standard_normal(shape)
    pad = np.zeros((n0, n1, 2 * nh))
    pad[:, :, :n2] = a
    buf = torch.tensor(pad, dtype=rdt, device=dev)
    plan = _plan(shape, prec, buf, buf)
    assert plan.inplace
    plan.execute_r2c(buf, buf)
    torch.cuda.synchronize()
    spec = buf.cpu().numpy().view(np.complex128 if prec == "f64" else np.complex64).reshape(n0, n1, nh)
    assert rel_l2(spec, orc.r2c(a)) < TOL[prec]
    plan.execute_c2r(buf, buf)
    torch.cuda.synchronize()
    back = buf.cpu().numpy()[:, :, :n2] / np.prod(shape)
    assert rel_l2(back, a) < TOL[prec]
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_preserve_input_flag(dev, orc, prec):
    import torch
    import accfft_amd as af
    rdt, cdt = _dt(prec)
    shape = (16, 8, 12)
    h = orc.r2c(rng(8).standard_normal(shape))
    hd = torch.tensor(h, dtype=cdt, device=dev)
    keep = hd.clone()
    out = torch.empty(shape, dtype=rdt, device=dev)
    plan = _plan(shape, prec, out, hd, flags=af.ACCFFT_MEASURE | af.PRESERVE_INPUT)
    plan.execute_c2r(hd, out)
    torch.cuda.synchronize()
    assert torch.equal(hd, keep)
    assert rel_l2(out.cpu().numpy(), orc.c2r(h, shape)) < TOL[prec]
    plan.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind", range(7))
def test_mult_wave_number(dev, orc, prec, kind):
    import torch
    import accfft_amd as af
    _, cdt = _dt(prec)
    N = (8, 7, 10)
    osize, ostart = (3, 7, 6), (4, 0, 0)   # a block at an x offset, like a pencil
    h = rng(9).standard_normal(osize) + 1j * rng(10).standard_normal(osize)
    hd = torch.tensor(h, dtype=cdt, device=dev)
    out = torch.empty_like(hd)
    mask = 0b011 if kind <= 2 else 0b111
    af.mult_wave_number(kind, out, hd, N, osize, ostart, mask)
    torch.cuda.synchronize()
    scale = 1.0 / (8 * 7 if kind <= 2 else 8 * 7 * 10)
    assert rel_l2(out.cpu().numpy(), orc.mult(kind, N, h, ostart, scale)) < TOL[prec]


def test_axpy(dev):
    import torch
    import accfft_amd as af
    x = torch.randn(1000, dtype=torch.float64, device=dev)
    y = torch.randn(1000, dtype=torch.float64, device=dev)
    ref = y + 2.5 * x
    af.axpy(2.5, x, y)
    torch.cuda.synchronize()
    # the kernel may contract to one FMA: agree to a few ulp
    assert torch.allclose(y, ref, rtol=1e-15, atol=1e-14)


@pytest.mark.parametrize("shape", [(512, 24, 40), (24, 1024, 34), (2048, 6, 10), (512, 8, 35)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_fused_two_level_operator_passes(dev, orc, shape, prec):
    """Strided axes of 512 / 1024 / 2048 points: the fused two-level tiles
    (forward FFT, symbol, inverse FFT in one pass) of FP32 plans for the
    spectral operators (x pass), grad / divergence (per-axis derivatives,
    complex view of column pairs; the odd z length takes the one-level
    tile) and the slow variants (y symbol pass); FP64 keeps the one-level
    tiles."""
    import torch
    import accfft_amd as af
    print(f"fused operator passes {shape} {prec}", flush=True)
    rdt, cdt, fp, tol = ((torch.float32, torch.complex64, af.F32, 1e-5) if prec == "f32"
                         else (torch.float64, torch.complex128, af.F64, 1e-12))
    a = rng(91).standard_normal(shape)
    ad = torch.tensor(a, dtype=rdt, device=dev)
    h = torch.empty((shape[0], shape[1], shape[2] // 2 + 1), dtype=cdt, device=dev)
    plan = af.Plan("r2c", shape, ad, h, fp)
    a64 = ad.cpu().numpy().astype(np.float64)
    out = torch.empty_like(ad)
    for name, ref in (("laplace", orc.laplace(a64)), ("inv_laplace", orc.inv_laplace(a64)),
                      ("biharmonic", orc.biharmonic(a64))):
        getattr(plan, name)(ad, out)
        torch.cuda.synchronize()
        assert rel_l2(out.cpu().numpy(), ref) < tol, name
    plan.poisson_solve(ad, out, 1)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), orc.poisson_solve(a64, 1)) < tol
    for slow in (False, True):
        g = [torch.empty_like(ad) for _ in range(3)]
        plan.grad(ad, *g, slow=slow)
        torch.cuda.synchronize()
        refs = orc.grad(a64, 7, slow)
        for d in range(3):
            assert rel_l2(g[d].cpu().numpy(), refs[d]) < tol, ("grad", slow, d)
        plan.divergence(g[0], g[1], g[2], out, slow=slow)
        torch.cuda.synchronize()
        gs = [x.cpu().numpy().astype(np.float64) for x in g]
        assert rel_l2(out.cpu().numpy(), orc.divergence(*gs, slow)) < tol, ("div", slow)
    plan.destroy()


def test_plan_on_two_streams(dev, orc):
    """One plan called from two streams: its shared workspace is ordered by
    the library (a call on a new stream waits for the previous stream's
    queued work), so both results match the oracle."""
    import torch
    import accfft_amd as af
    n = (64, 48, 40)
    u = [rng(200 + i).standard_normal(n) for i in range(2)]
    ud = [torch.tensor(a, device=dev) for a in u]
    out = [torch.empty_like(ud[0]) for _ in range(2)]
    plan = af.Plan("r2c", n, ud[0], None, af.F64)
    s = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    for rep in range(3):
        for i in range(2):
            with torch.cuda.stream(s[i]):
                plan.laplace(ud[i], out[i])
                plan.inv_laplace(out[i], out[i])
        torch.cuda.synchronize()
        for i in range(2):
            ref = orc.inv_laplace(orc.laplace(u[i]))
            assert rel_l2(out[i].cpu().numpy(), ref) < 1e-12, (rep, i)
    plan.destroy()
