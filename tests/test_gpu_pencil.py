"""GPU pencil exchange copies (accfft_amd_pencil_pack / _unpack, the
memcpy_v1_h1 / memcpy_v1_h2 of src/transpose_cuda.cu:97-197): bit-exact
against the oracle on ragged splits, empty ranks and several element
sizes, plus the P-rank pencil transpose run through the GPU kernels."""
import numpy as np
import pytest

from conftest import rng
from test_pencil_oracle import _split, simulate_pencil_transpose

pytestmark = pytest.mark.gpu

DTYPES = ["float32", "float64", "complex64", "complex128"]


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("howmany,n0,N1,t,P,empty", [(1, 4, 10, 3, 3, None), (3, 5, 7, 2, 4, 2),
                                                     (2, 64, 300, 130, 8, None), (1, 1, 1, 1, 1, None)])
def test_pack_matches_oracle(dev, orc, dtype, howmany, n0, N1, t, P, empty):
    import torch
    import accfft_amd as af
    g = rng(71)
    a = g.standard_normal((howmany, n0, N1, t)).astype(dtype)
    if np.iscomplexobj(a):
        a = a + 1j * g.standard_normal(a.shape).astype(a.real.dtype)
    sizes, starts = _split(N1, P, empty)
    ref = orc.pencil_pack(P, howmany, n0, t, sizes, a, n0 * N1 * t, N1, starts)
    ad = _t(a, dev)
    out = torch.zeros(ref.size, dtype=ad.dtype, device=dev)
    af.memcpy_v1_h1(P, howmany, n0, t, sizes, out, ad, n0 * N1 * t, N1, starts)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("howmany,N0,n1,t,P,empty", [(1, 10, 4, 3, 3, None), (3, 7, 5, 2, 4, 0),
                                                     (2, 300, 64, 130, 8, None)])
def test_unpack_matches_oracle(dev, orc, dtype, howmany, N0, n1, t, P, empty):
    import torch
    import accfft_amd as af
    sizes, starts = _split(N0, P, empty)
    recv = rng(72).standard_normal(howmany * N0 * n1 * t).astype(dtype)
    base = np.full(howmany * N0 * n1 * t, -7, dtype=dtype)
    ref = orc.pencil_unpack(P, howmany, starts, sizes, base, N0 * n1 * t, n1, t, recv)
    d = _t(base, dev)
    af.memcpy_v1_h2(P, howmany, starts, sizes, d, N0 * n1 * t, n1, t, _t(recv, dev))
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), ref)


@pytest.mark.parametrize("shape,P,empty", [((2, 12, 10, 3), 3, None), ((3, 256, 200, 9), 8, 3)])
def test_pencil_transpose_on_gpu(dev, shape, P, empty):
    import torch
    import accfft_amd as af

    def pack(P, howmany, n0, t, cols, loc, idist, N1, cstart):
        ld = _t(loc, dev)
        out = torch.empty(ld.numel(), dtype=ld.dtype, device=dev)
        af.memcpy_v1_h1(P, howmany, n0, t, cols, out, ld, idist, N1, cstart)
        torch.cuda.synchronize()
        return out.cpu().numpy()

    def unpack(P, howmany, rstart, rows, data, odist, n1, t, recv):
        d = _t(data, dev)
        af.memcpy_v1_h2(P, howmany, rstart, rows, d, odist, n1, t, _t(recv, dev))
        torch.cuda.synchronize()
        return d.cpu().numpy()

    G = rng(73).standard_normal(shape)
    for res, cs, cn in simulate_pencil_transpose(pack, unpack, G, P, empty):
        assert np.array_equal(res, G[:, :, cs:cs + cn, :].ravel())


def test_pencil_rejects_too_many_ranks():
    import accfft_amd as af
    with pytest.raises(af.AccfftError):
        af.lib()
        import torch
        x = torch.zeros(4)
        af.memcpy_v1_h1(65, 1, 1, 1, [0] * 65, x, x, 1, 1, [0] * 65)
