"""accfft_amd_inv_laplace_noalloc (no counterpart in the reference): the
inverse Laplacian with its intermediate spectrum in the library's fixed
device scratch instead of the plan's workspace.  Called through the C ABI
and checked against accfft_amd_inv_laplace (same kernels: bit-identical)
and the FP64 oracle, on two calls in a row -- a 256^3 FP32 grid, then a
smaller FP64 one -- and the plan never allocates a workspace."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _noalloc(L):
    f = L.accfft_amd_inv_laplace_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def _case(dev, orc, n, dtype, seed):
    import torch
    import accfft_amd as af
    L = af.lib()
    fp = af.F32 if dtype == torch.float32 else af.F64
    print(f"no-alloc inv_laplace {n} {dtype}", flush=True)
    u0 = np.random.default_rng(seed).standard_normal(n)
    u = torch.tensor(u0, dtype=dtype, device=dev)
    ref = torch.empty_like(u)
    out = torch.full_like(u, float("nan"))
    # the allocating entry point on one plan, the scratch one on a fresh plan
    p1 = af.Plan("r2c", n, u, None, fp)
    p1.inv_laplace(u, ref)
    p2 = af.Plan("r2c", n, u, None, fp)
    st = _noalloc(L)(p2.handle, out.data_ptr(), u.data_ptr(), None)
    assert st == 0, L.accfft_amd_status_string(st)
    torch.cuda.synchronize()
    assert p2.workspace_bytes() == 0  # the plan allocated nothing
    assert torch.equal(out, ref)
    tol = 1e-5 if dtype == torch.float32 else 1e-12
    assert rel_l2(out.cpu().numpy(), orc.inv_laplace(u.cpu().numpy().astype(np.float64))) < tol
    p1.destroy()
    p2.destroy()


def test_noalloc_matches_allocating_entry_point(dev, orc):
    import torch
    _case(dev, orc, (256, 256, 256), torch.float32, 1)
    # second call, smaller input: the scratch keeps the first call's contents
    _case(dev, orc, (24, 20, 18), torch.float64, 2)


def test_noalloc_rejects_oversize(dev):
    import torch
    import accfft_amd as af
    L = af.lib()
    n = (512, 512, 256)  # a 270 MB FP32 spectrum: above the 160 MiB scratch
    u = torch.zeros(1, device=dev)
    p = af.Plan("r2c", n, None, None, af.F32, af.ACCFFT_ESTIMATE)
    st = _noalloc(L)(p.handle, u.data_ptr(), u.data_ptr(), None)
    assert st == 8, st  # ACCFFT_AMD_ERR_SCRATCH, before any launch
    assert b"scratch" in L.accfft_amd_status_string(st)
    p.destroy()
    # an axis longer than one LDS tile needs the plan's own workspaces
    q = af.Plan("r2c", (2, 3, 16384), None, None, af.F32, af.ACCFFT_ESTIMATE)
    st = _noalloc(L)(q.handle, u.data_ptr(), u.data_ptr(), None)
    assert st == 3, st  # ACCFFT_AMD_ERR_UNSUPPORTED, before any launch
    q.destroy()
