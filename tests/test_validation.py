"""The Python binding refuses tensors the kernels cannot use (host memory,
non-contiguous views, the wrong precision, too few bytes) before any call
reaches the C ABI, so a bad argument raises instead of making the kernels
read or write out of bounds."""
import pytest


def test_arg_checks_on_cpu():
    import torch
    import accfft_amd as af
    with pytest.raises(af.AccfftError, match="not a GPU"):
        af._arg(torch.zeros(8), 0, af.F32, "x")
    with pytest.raises(af.AccfftError, match="required"):
        af._arg(None, 0, af.F32, "x", optional=False)
    assert af._arg(None, 0, af.F32, "x") is None
    with pytest.raises(af.AccfftError, match="torch.Tensor"):
        af._arg([1.0], 0, af.F32, "x")


@pytest.mark.gpu
def test_plan_rejects_bad_tensors(dev):
    import torch
    import accfft_amd as af
    n = (16, 12, 10)
    x = torch.randn(n, device=dev)
    h = torch.empty((16, 12, 6), dtype=torch.complex64, device=dev)
    plan = af.Plan("r2c", n, x, h, af.F32)
    plan.execute_r2c(x, h)  # the good call
    with pytest.raises(af.AccfftError, match="not contiguous"):
        plan.execute_r2c(x.transpose(0, 1), h)
    with pytest.raises(af.AccfftError, match="not contiguous"):
        plan.execute_c2r(h, torch.empty((16, 12, 20), device=dev)[:, :, ::2])
    with pytest.raises(af.AccfftError, match="precision"):
        plan.execute_r2c(x.double(), h)
    with pytest.raises(af.AccfftError, match="bytes"):
        plan.execute_r2c(x[:8], h)
    with pytest.raises(af.AccfftError, match="bytes"):
        plan.execute_r2c(x, h[:, :, :3].contiguous())
    with pytest.raises(af.AccfftError, match="not a GPU"):
        plan.execute_r2c(x.cpu(), h)
    with pytest.raises(af.AccfftError, match="bytes"):
        plan.laplace(x, torch.empty(10, device=dev))
    with pytest.raises(af.AccfftError, match="bytes"):
        plan.execute_r2c_axis(1, x, torch.empty((16, 6, 10), dtype=torch.complex64, device=dev))
    plan.execute_r2c_axis(1, x, torch.empty((16, 7, 10), dtype=torch.complex64, device=dev))
    torch.cuda.synchronize()
    plan.destroy()
