"""GPU parity against the FP64 oracle at the workloads' full sizes, on the
workloads' own seeded inputs (accfft_amd/fields.py):

- the plain method at 64^3 FP64: sin x + sin y + sin z + N(0, 1e-3) noise;
  spectrum and round trip (src/accfft_gpu.cpp:444-623,
  steps/step1/step1_gpu.cpp:52-84);
- the drop-in at 256^3 and 512^3, FP64 and FP32, on i.i.d. N(0, 1) and on
  the band-limited |k|^(-5/3) random-phase field: the full r2c spectrum
  (every bin, so a permuted or misplaced bin fails), the c2r of the
  oracle's spectrum, grad with its three outputs, laplace and inverse
  laplace (src/operators_gpu.txx:172-334, src/operators.txx:1270-1321);
- the precision forms at 512^3 FP32 (r2c and c2c): the FP32 VALU path and
  the split BF16x2 / BF16x3 / FP16x2 matrix-core stages on the two inputs
  above plus N(0, 1) * 10^U(-4, 4); relative L2 and max-abs error against
  the FP64 oracle are printed for each.

Bounds: relative L2 <= 1e-12 (FP64) and <= 1e-5 (FP32 and every split
form).  Errors are computed on the GPU against the oracle's output copied
there in FP64."""
import numpy as np
import pytest

from accfft_amd import fields

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}
SPLIT_FLAGS = {"fp32": 0, "bf16x2": 0x1000, "bf16x3": 0x2000, "fp16x2": 0x4000}


def _dt(prec, cplx=False):
    import torch
    if cplx:
        return torch.complex128 if prec == "f64" else torch.complex64
    return torch.float64 if prec == "f64" else torch.float32


def _errs(out, ref_np, dev):
    """(relative L2, max |out - ref|, max |ref|) with the reference in FP64."""
    import torch
    ref = torch.from_numpy(np.ascontiguousarray(ref_np)).to(dev)
    o = out.to(ref.dtype)
    d = o - ref
    nr = torch.linalg.vector_norm(ref).item()
    rel = torch.linalg.vector_norm(d).item() / nr if nr > 0 else torch.linalg.vector_norm(d).item()
    mx = d.abs().max().item()
    mref = ref.abs().max().item()
    del ref, o, d
    return rel, mx, mref


def test_plain_method_64_fp64(dev, orc):
    """Workload 0: 64^3 FP64 r2c then c2r of sin x + sin y + sin z + N(0, 1e-3)."""
    import torch
    import accfft_amd as af
    n = (64, 64, 64)
    x = fields.sines_noise(n, fields.SEED0 + 0)
    print(f"plain method 64^3 checksum {fields.checksum(x)}", flush=True)
    xd = torch.tensor(x, device=dev)
    h = torch.empty((64, 64, 33), dtype=torch.complex128, device=dev)
    back = torch.empty_like(xd)
    plan = af.Plan("r2c", n, xd, h, af.F64)
    plan.execute_r2c(xd, h)
    plan.execute_c2r(h.clone(), back)
    torch.cuda.synchronize()
    ref = orc.r2c(x)
    rel, mx, mref = _errs(h, ref, dev)
    print(f"  r2c rel-L2 {rel:.3e} max-abs {mx:.3e} (max |X| {mref:.3e})")
    assert rel < 1e-12
    # the noise-free bins of the closed form: -i N^3/2 at (1,0,0), (0,1,0), (0,0,1)
    hc = h.cpu().numpy()
    for idx in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        assert abs(hc[idx] - ref[idx]) < 1e-12 * 64 ** 3
        assert abs(hc[idx] + 0.5j * 64 ** 3) < 64 ** 3 * 1e-4  # noise of sigma 1e-3
    rt = np.linalg.norm(back.cpu().numpy() / 64 ** 3 - x) / np.linalg.norm(x)
    print(f"  round trip rel-L2 {rt:.3e}")
    assert rt < 1e-12
    assert _errs(back, orc.c2r(ref, n), dev)[0] < 1e-12
    plan.destroy()


_DIST = {"normal": lambda n, s: fields.normal(n, s),
         "band_limited": lambda n, s: fields.band_limited(n, s, slope=-5.0 / 3.0)}


@pytest.mark.parametrize("N", [256, 512])
@pytest.mark.parametrize("dist", ["normal", "band_limited"])
def test_dropin_full_size(dev, orc, N, dist):
    """Workload 1 at its full sizes: r2c spectrum, c2r, grad (3 outputs),
    laplace and inverse laplace, FP64 and FP32, against the FP64 oracle."""
    import torch
    import accfft_amd as af
    n = (N, N, N)
    x = _DIST[dist](n, fields.SEED0 + 1)
    print(f"drop-in {n} {dist} checksum {fields.checksum(x)}", flush=True)
    ref_h = orc.r2c(x)
    ref_b = orc.c2r(ref_h, n)
    for prec in ("f64", "f32"):
        fp = af.F64 if prec == "f64" else af.F32
        xd = torch.tensor(x, dtype=_dt(prec), device=dev)
        h = torch.empty((N, N, N // 2 + 1), dtype=_dt(prec, True), device=dev)
        plan = af.Plan("r2c", n, xd, h, fp)
        plan.execute_r2c(xd, h)
        torch.cuda.synchronize()
        rel, mx, mref = _errs(h, ref_h, dev)
        print(f"  {prec} r2c rel-L2 {rel:.3e} max-abs {mx:.3e} (max |X| {mref:.3e})", flush=True)
        assert rel < TOL[prec]
        del h
        # c2r of the oracle's spectrum
        hd = torch.from_numpy(ref_h).to(dev).to(_dt(prec, True))
        back = torch.empty_like(xd)
        plan.execute_c2r(hd, back)
        torch.cuda.synchronize()
        del hd
        rel = _errs(back, ref_b, dev)[0]
        print(f"  {prec} c2r rel-L2 {rel:.3e}", flush=True)
        assert rel < TOL[prec]
        plan.destroy()
        del back
        torch.cuda.empty_cache()
    del ref_h, ref_b
    # operators: the oracle's outputs one at a time
    ops = [("grad", lambda: orc.grad(x, 7)), ("laplace", lambda: [orc.laplace(x)]),
           ("inv_laplace", lambda: [orc.inv_laplace(x)])]
    for name, make_ref in ops:
        ref = make_ref()
        for prec in ("f64", "f32"):
            fp = af.F64 if prec == "f64" else af.F32
            xd = torch.tensor(x, dtype=_dt(prec), device=dev)
            h = torch.empty((N, N, N // 2 + 1), dtype=_dt(prec, True), device=dev)
            plan = af.Plan("r2c", n, xd, h, fp)
            outs = [torch.empty_like(xd) for _ in ref]
            if name == "grad":
                plan.grad(xd, *outs)
            else:
                getattr(plan, name)(xd, outs[0])
            torch.cuda.synchronize()
            for c, (o, r) in enumerate(zip(outs, ref)):
                rel, mx, mref = _errs(o, r, dev)
                print(f"  {prec} {name}[{c}] rel-L2 {rel:.3e} max-abs {mx:.3e} "
                      f"(max {mref:.3e})", flush=True)
                assert rel < TOL[prec], (name, c, prec)
            plan.destroy()
            del outs, xd, h
            torch.cuda.empty_cache()
        del ref


_DIST3 = {"normal": fields.normal,
          "band_limited": lambda n, s: fields.band_limited(n, s, slope=-5.0 / 3.0),
          "dynamic_range": fields.dynamic_range}


@pytest.mark.parametrize("dist", ["normal", "band_limited", "dynamic_range"])
def test_precision_forms_512(dev, orc, dist):
    """Workload 2: 512^3 FP32 r2c and c2c, the FP32 VALU path and the split
    BF16x2 / BF16x3 / FP16x2 forms, against the FP64 oracle."""
    import torch
    import accfft_amd as af
    n = (512, 512, 512)
    s = fields.SEED0 + 2
    x = _DIST3[dist](n, s)
    print(f"precision forms 512^3 {dist} r2c", flush=True)
    # every form sees the same FP32 input; the oracle transforms that input
    x32 = x.astype(np.float32)
    del x
    ref = orc.r2c(x32.astype(np.float64))
    xd = torch.from_numpy(x32).to(dev)
    h = torch.empty((512, 512, 257), dtype=torch.complex64, device=dev)
    worst = {}
    for form, flag in SPLIT_FLAGS.items():
        plan = af.Plan("r2c", n, xd, h, af.F32, af.ACCFFT_MEASURE | flag)
        plan.execute_r2c(xd, h)
        torch.cuda.synchronize()
        rel, mx, mref = _errs(h, ref, dev)
        print(f"  r2c {form:7s} rel-L2 {rel:.3e} max-abs {mx:.3e} max-abs/max|X| {mx / mref:.3e}",
              flush=True)
        worst[("r2c", form)] = rel
        plan.destroy()
    del ref, xd, h, x32
    torch.cuda.empty_cache()
    print(f"precision forms 512^3 {dist} c2c", flush=True)
    c = fields.complex_of(_DIST3[dist], n, s).astype(np.complex64)
    ref = orc.c2c(c.astype(np.complex128), -1)
    cd = torch.from_numpy(c).to(dev)
    del c
    out = torch.empty_like(cd)
    for form, flag in SPLIT_FLAGS.items():
        plan = af.Plan("c2c", n, cd, out, af.F32, af.ACCFFT_MEASURE | flag)
        plan.execute_c2c(af.ACCFFT_FORWARD, cd, out)
        torch.cuda.synchronize()
        rel, mx, mref = _errs(out, ref, dev)
        print(f"  c2c {form:7s} rel-L2 {rel:.3e} max-abs {mx:.3e} max-abs/max|X| {mx / mref:.3e}",
              flush=True)
        worst[("c2c", form)] = rel
        plan.destroy()
    for k, v in worst.items():
        assert v < 1e-5, (k, v)
