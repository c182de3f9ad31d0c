"""ACCFFT_MEASURE tuning of the strided tile configurations (doc/flags.md;
include/accfft_common.h:40-42).  The tuner may pick any configuration, so
every one of them is checked against the FP64 oracle (rel-L2 <= 1e-5 FP32,
<= 1e-12 FP64), on plain rows and on the slab plan's packed rows; MEASURE
must report a valid choice, ESTIMATE the built-in rule."""
import numpy as np
import pytest

from conftest import rel_l2, rng

pytestmark = pytest.mark.gpu

CANDIDATES = [-2, 0, 1, 2, 3]


def _choices(plan, axis):
    out = []
    for c in CANDIDATES:
        try:
            plan.set_tuning(axis, c)
            out.append(c)
        except Exception:
            pass
    plan.set_tuning(axis, -1)
    return out


@pytest.mark.parametrize("shape,prec", [((512, 512, 30), "f32"), ((1024, 6, 20), "f32"),
                                        ((2048, 4, 16), "f32"), ((6, 2048, 10), "f32"),
                                        ((512, 20, 12), "f64")])
def test_every_configuration_matches_oracle(dev, orc, shape, prec):
    import torch
    import accfft_amd as af
    rdt, cdt, fp, tol = ((torch.float32, torch.complex64, af.F32, 1e-5) if prec == "f32"
                         else (torch.float64, torch.complex128, af.F64, 1e-12))
    a = torch.tensor(rng(81).standard_normal(shape), dtype=rdt, device=dev)
    h = torch.empty((shape[0], shape[1], shape[2] // 2 + 1), dtype=cdt, device=dev)
    back = torch.empty_like(a)
    plan = af.Plan("r2c", shape, a, h, fp, af.ACCFFT_ESTIMATE)
    ref = orc.r2c(a.cpu().numpy().astype(np.float64))
    for axis in (0, 1):
        if shape[axis] not in (512, 1024, 2048):
            continue
        for c in _choices(plan, axis):
            print(f"{shape} {prec} axis {axis} choice {c}", flush=True)
            plan.set_tuning(axis, c)
            plan.execute_r2c(a, h)
            torch.cuda.synchronize()
            assert rel_l2(h.cpu().numpy(), ref) < tol, (axis, c)
            plan.execute_c2r(h, back)
            torch.cuda.synchronize()
            e = (torch.linalg.vector_norm(back / float(np.prod(shape)) - a) /
                 torch.linalg.vector_norm(a)).item()
            assert e < 2 * tol, (axis, c, e)
        plan.set_tuning(axis, -1)
    plan.destroy()


@pytest.mark.parametrize("N,P", [((64, 512, 40), 4), ((1024, 1024, 16), 8)])
def test_slab_configurations_match_oracle(dev, orc, N, P):
    import torch
    import accfft_amd as af
    n0, n1, Nh = N[0] // P, N[1] // P, N[2] // 2 + 1
    u = torch.tensor(rng(82).standard_normal((n0, N[1], N[2])), dtype=torch.float32, device=dev)
    h = torch.empty((N[0], n1, Nh), dtype=torch.complex64, device=dev)
    plan = af.Plan("r2c_slab", N, u, h, af.F32, af.ACCFFT_ESTIMATE, nprocs=P)
    ref = orc.slab_r2c(u.cpu().numpy().astype(np.float64), P)
    for axis in (0, 1):
        if N[axis] not in (512, 1024, 2048):
            continue
        for c in _choices(plan, axis):
            print(f"slab {N} P {P} axis {axis} choice {c}", flush=True)
            plan.set_tuning(axis, c)
            plan.execute_r2c(u, h)
            torch.cuda.synchronize()
            assert rel_l2(h.cpu().numpy(), ref) < 1e-5, (axis, c)
        plan.set_tuning(axis, -1)
    plan.destroy()


def test_measure_reports_a_valid_choice(dev):
    import torch
    import accfft_amd as af
    n = (512, 512, 512)
    x = torch.randn(n, device=dev)
    h = torch.empty((512, 512, 257), dtype=torch.complex64, device=dev)
    pm = af.Plan("r2c", n, x, h, af.F32, af.ACCFFT_MEASURE)
    pe = af.Plan("r2c", n, x, h, af.F32, af.ACCFFT_ESTIMATE)
    tm, te = pm.tuning(), pe.tuning()
    print(f"512^3 FP32 MEASURE choice {tm}")
    assert te == (-1, -1, -1)
    assert all(c in (-1, -2, 0, 1, 2, 3) for c in tm) and tm[2] == -1
    pm.destroy()
    pe.destroy()


def test_set_tuning_rejects_missing_variants(dev):
    import torch
    import accfft_amd as af
    x = torch.randn((64, 48, 40), device=dev)
    h = torch.empty((64, 48, 21), dtype=torch.complex64, device=dev)
    plan = af.Plan("r2c", (64, 48, 40), x, h, af.F32, af.ACCFFT_ESTIMATE)
    with pytest.raises(af.AccfftError):
        plan.set_tuning(0, 0)  # 64 has no two-level configuration
    plan.set_tuning(0, -2)
    with pytest.raises(af.AccfftError):
        plan.set_tuning(3, -1)
    plan.destroy()
