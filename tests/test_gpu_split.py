"""Split-precision matrix-core radix-16 stages (FP32 plans with
ACCFFT_AMD_SPLIT_BF16 / _BF16X3 / _FP16).

The one-wavefront DFT-16 self-test pins the v_mfma_f32_32x32x16 lane layout
(a wrong row/column/accumulator mapping gives O(1) errors) and the split
error of one stage.  The 3-D transforms are then compared with the FP64
oracle; the measured error of each split form is printed and held to the
drop-in FP32 bound, relative L2 <= 1e-5 (one matrix-core stage per axis:
per stage ~2^-17 for 2-term BF16, ~2^-22 for 2-term FP16 and 3-term BF16).
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

SPLITS = {"bf16x2": (1, 0x1000), "bf16x3": (2, 0x2000), "fp16x2": (3, 0x4000)}
# bound on rel-L2 vs FP64 for a full 3-D transform
BOUND3D = {"bf16x2": 1e-5, "bf16x3": 1e-5, "fp16x2": 1e-5}


@pytest.mark.parametrize("split", list(SPLITS))
@pytest.mark.parametrize("sign", [-1, 1])
def test_dft16_selftest(dev, split, sign):
    import torch
    import accfft_amd as af
    kind = SPLITS[split][0]
    x = (rng(5).standard_normal((32, 16)) + 1j * rng(6).standard_normal((32, 16)))
    # one vector with a huge dynamic range, one all-zero vector
    x[3] *= 10.0 ** rng(7).uniform(-4, 4, 16)
    x[7] = 0
    xd = torch.tensor(x, dtype=torch.complex64, device=dev)
    yd = torch.zeros_like(xd)
    L = af.lib()
    L.accfft_amd_selftest_dft16.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]
    assert L.accfft_amd_selftest_dft16(kind, sign, xd.data_ptr(), yd.data_ptr(), None) == 0
    torch.cuda.synchronize()
    xr = xd.cpu().numpy().astype(np.complex128)
    ref = np.fft.fft(xr, axis=1) if sign < 0 else np.fft.ifft(xr, axis=1) * 16
    y = yd.cpu().numpy()
    err = rel_l2(y, ref)
    print(f"dft16 {split} sign {sign}: rel-L2 {err:.3e}")
    assert err < (1e-5 if split == "bf16x2" else 1e-6)
    # every vector individually (layout: each output lands in its own slot)
    for b in range(32):
        nb = np.linalg.norm(ref[b])
        if nb > 0:
            assert np.linalg.norm(y[b] - ref[b]) / nb < (1e-4 if split == "bf16x2" else 1e-5), b
        else:
            assert np.all(y[b] == 0)


def _split_plan(kind, n, a, b, split):
    import accfft_amd as af
    return af.Plan(kind, n, a, b, af.F32, af.ACCFFT_MEASURE | SPLITS[split][1])


@pytest.mark.parametrize("split", list(SPLITS))
@pytest.mark.parametrize("shape", [(256, 256, 256), (64, 512, 1024), (2048, 16, 256), (16, 4096, 8),
                                   (8, 32, 512), (4, 8, 2048), (2, 4, 4096)])
def test_split_r2c_c2r(dev, orc, split, shape):
    import torch
    print(f"split {split} r2c/c2r {shape}", flush=True)
    a = rng(11).standard_normal(shape)
    ad = torch.tensor(a, dtype=torch.float32, device=dev)
    h = torch.empty((shape[0], shape[1], shape[2] // 2 + 1), dtype=torch.complex64, device=dev)
    plan = _split_plan("r2c", shape, ad, h, split)
    plan.execute_r2c(ad, h)
    torch.cuda.synchronize()
    ref = orc.r2c(ad.cpu().numpy().astype(np.float64))
    e1 = rel_l2(h.cpu().numpy(), ref)
    back = torch.empty_like(ad)
    plan.execute_c2r(h, back)
    torch.cuda.synchronize()
    e2 = rel_l2(back.cpu().numpy() / np.prod(shape), ad.cpu().numpy())
    print(f"  r2c rel-L2 {e1:.3e}, round trip {e2:.3e}")
    assert e1 < BOUND3D[split] and e2 < 2 * BOUND3D[split]
    plan.destroy()


@pytest.mark.parametrize("split", list(SPLITS))
def test_split_c2c_dynamic_range(dev, orc, split):
    """N(0,1) * 10^U(-4,4) per point: the large-dynamic-range input."""
    import torch
    shape = (256, 64, 512)
    g = rng(12)
    c = (g.standard_normal(shape) + 1j * g.standard_normal(shape)) * 10.0 ** g.uniform(-4, 4, shape)
    cd = torch.tensor(c, dtype=torch.complex64, device=dev)
    out = torch.empty_like(cd)
    plan = _split_plan("c2c", shape, cd, out, split)
    plan.execute_c2c(-1, cd, out)
    torch.cuda.synchronize()
    ref = orc.c2c(cd.cpu().numpy().astype(np.complex128), -1)
    err = rel_l2(out.cpu().numpy(), ref)
    print(f"split {split} c2c dynamic-range rel-L2 {err:.3e}")
    assert err < BOUND3D[split]
    plan.destroy()


def test_split_rejected_for_fp64():
    import accfft_amd as af
    with pytest.raises(af.AccfftError):
        af.Plan("r2c", (256, 256, 256), None, None, af.F64, af.ACCFFT_MEASURE | 0x1000)


@pytest.mark.parametrize("split", list(SPLITS))
@pytest.mark.parametrize("shape", [(1024, 8, 64), (2048, 4, 32), (8, 512, 48), (512, 6, 40)])
@pytest.mark.parametrize("direction", [-1, 1])
def test_split_two_level_c2c(dev, orc, split, shape, direction):
    """Strided axes of 512, 1024 and 2048 points: the two-level tiles whose
    first radix-16 level runs on the matrix cores (odd column counts leave
    a partial last tile)."""
    import torch
    print(f"split {split} two-level c2c {shape} dir {direction}", flush=True)
    g = rng(13)
    c = g.standard_normal(shape) + 1j * g.standard_normal(shape)
    cd = torch.tensor(c, dtype=torch.complex64, device=dev)
    out = torch.empty_like(cd)
    plan = _split_plan("c2c", shape, cd, out, split)
    plan.execute_c2c(direction, cd.clone(), out)
    torch.cuda.synchronize()
    ref = orc.c2c(cd.cpu().numpy().astype(np.complex128), direction)
    err = rel_l2(out.cpu().numpy(), ref)
    print(f"  rel-L2 {err:.3e}")
    assert err < BOUND3D[split]
    plan.destroy()


@pytest.mark.parametrize("split", list(SPLITS))
@pytest.mark.parametrize("shape", [(512, 512, 30), (1024, 6, 20), (2048, 4, 16)])
def test_split_every_configuration(dev, orc, split, shape):
    """Every tile configuration the MEASURE tuning may keep for a split
    plan's strided axes (one-level, the two-level variants, persistent
    tiles included) against the FP64 oracle, forward and round trip."""
    import torch
    import accfft_amd as af
    a = torch.tensor(rng(83).standard_normal(shape), dtype=torch.float32, device=dev)
    h = torch.empty((shape[0], shape[1], shape[2] // 2 + 1), dtype=torch.complex64, device=dev)
    back = torch.empty_like(a)
    plan = af.Plan("r2c", shape, a, h, af.F32, af.ACCFFT_ESTIMATE | SPLITS[split][1])
    ref = orc.r2c(a.cpu().numpy().astype(np.float64))
    for axis in (0, 1):
        if shape[axis] not in (512, 1024, 2048):
            continue
        for c in (-2, 0, 1, 2, 3):
            try:
                plan.set_tuning(axis, c)
            except Exception:
                continue
            print(f"split {split} {shape} axis {axis} choice {c}", flush=True)
            plan.execute_r2c(a, h)
            torch.cuda.synchronize()
            err = rel_l2(h.cpu().numpy(), ref)
            assert err < BOUND3D[split], (axis, c, err)
            plan.execute_c2r(h, back)
            torch.cuda.synchronize()
            e = (torch.linalg.vector_norm(back / float(np.prod(shape)) - a) /
                 torch.linalg.vector_norm(a)).item()
            assert e < 2 * BOUND3D[split], (axis, c, e)
        plan.set_tuning(axis, -1)
    plan.destroy()
