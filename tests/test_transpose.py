"""Local transposes of src/transpose_cuda.cu (and src/transpose.cpp:664-769).

CPU: the oracle's permutations against independent numpy / pure-Python
brute force on tiny shapes, ragged tuples included.
GPU: the HIP kernels against the oracle, bit-exact, for float, double and
complex elements, plus the pack/unpack shapes of a 2048^3 slab split over 8
ranks (local_n0 = local_n1 = 256, t = 2 * 1025 reals) at reduced t.
"""
import numpy as np
import pytest

from conftest import rng


def brute_ragged(r, c, t, t2, a):
    """Unpack: row r-1 of the input holds c tuples of t2 elements."""
    out = []
    for j in range(c):
        for i in range(r):
            if i == r - 1:
                s = i * c * t + j * t2
                out.append(a[s:s + t2])
            else:
                s = (i * c + j) * t
                out.append(a[s:s + t])
    return np.concatenate(out) if out else a[:0]


def brute_col(r, c, t, t2, a):
    """Pack: column c-1 of every input row holds a tuple of t2 elements."""
    row = (c - 1) * t + t2
    out = []
    for j in range(c):
        for i in range(r):
            ln = t2 if j == c - 1 else t
            s = i * row + j * t
            out.append(a[s:s + ln])
    return np.concatenate(out)


CASES = [(1, 1, 1), (2, 3, 1), (3, 2, 4), (5, 7, 3), (8, 8, 2), (4, 9, 5), (16, 1, 3), (1, 16, 3)]


@pytest.mark.parametrize("r,c,t", CASES)
def test_oracle_plain(orc, r, c, t):
    a = rng(r * 100 + c).standard_normal(r * c * t)
    ref = a.reshape(r, c, t).transpose(1, 0, 2).ravel()
    np.testing.assert_array_equal(orc.local_transpose(r, c, t, a), ref)


@pytest.mark.parametrize("r,c,t", CASES)
@pytest.mark.parametrize("t2", [1, 2])
def test_oracle_ragged(orc, r, c, t, t2):
    t2 = min(t2, t)
    n = c * ((r - 1) * t + t2)
    a = rng(7).standard_normal(n)
    np.testing.assert_array_equal(orc.local_transpose_ragged(r, c, t, t2, a),
                                  brute_ragged(r, c, t, t2, a))
    m = r * ((c - 1) * t + t2)
    b = rng(8).standard_normal(m)
    np.testing.assert_array_equal(orc.local_transpose_col(r, c, t, t2, b),
                                  brute_col(r, c, t, t2, b))


def test_pack_unpack_roundtrip(orc):
    """Pack (local_transpose_col) of [local_n0][P][local_n1*t] followed by the
    unpack (local_transpose) restores the x-major layout with an identity
    exchange (transpose_gpu.cpp:3889-3891, :1411-1413)."""
    P, ln, t = 4, 3, 10
    a = rng(9).standard_normal(ln * P * ln * t)
    packed = orc.local_transpose_col(ln, P, ln * t, ln * t, a)
    assert packed.size == a.size
    back = orc.local_transpose(P, ln, ln * t, packed)
    np.testing.assert_array_equal(back, a)


# ------------------------------------------------------------------ GPU
DT = {"f32": np.float32, "f64": np.float64, "c64": np.complex64}


@pytest.mark.gpu
@pytest.mark.parametrize("dt", ["f32", "f64", "c64"])
@pytest.mark.parametrize("r,c,t,t2", [(3, 5, 7, 3), (8, 8, 2, 2), (4, 9, 1000, 999), (7, 3, 4097, 5),
                                      (256, 8, 3, 3), (1, 1, 1, 1)])
def test_gpu_transposes(dev, orc, dt, r, c, t, t2):
    import torch
    import accfft_amd as af
    print(f"transpose {dt} r={r} c={c} t={t} t2={t2}", flush=True)
    g = rng(r + c + t)
    mk = (lambda n: (g.standard_normal(n) + 1j * g.standard_normal(n)).astype(DT[dt])) \
        if dt == "c64" else (lambda n: g.standard_normal(n).astype(DT[dt]))
    tt = {"f32": torch.float32, "f64": torch.float64, "c64": torch.complex64}[dt]
    # plain
    a = mk(r * c * t)
    ad = torch.tensor(a, dtype=tt, device=dev)
    od = torch.empty_like(ad)
    af.local_transpose(r, c, t, ad, od)
    # in place
    ip = ad.clone()
    af.local_transpose_inplace(r, c, t, ip)
    # ragged unpack
    ar = mk(c * ((r - 1) * t + t2))
    ard = torch.tensor(ar, dtype=tt, device=dev)
    ord_ = torch.empty_like(ard)
    af.local_transpose_ragged(r, c, t, t2, ard, ord_)
    # ragged pack
    ac = mk(r * ((c - 1) * t + t2))
    acd = torch.tensor(ac, dtype=tt, device=dev)
    ocd = torch.empty_like(acd)
    af.local_transpose_col(r, c, t, t2, acd, ocd)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(od.cpu().numpy(), orc.local_transpose(r, c, t, a))
    np.testing.assert_array_equal(ip.cpu().numpy(), orc.local_transpose(r, c, t, a))
    np.testing.assert_array_equal(ord_.cpu().numpy(), orc.local_transpose_ragged(r, c, t, t2, ar))
    np.testing.assert_array_equal(ocd.cpu().numpy(), orc.local_transpose_col(r, c, t, t2, ac))


@pytest.mark.gpu
def test_gpu_slab_pack_unpack(dev, orc):
    """Shapes of one rank of a 2048^3 FP32 slab split over 8 ranks
    (local_n0 = local_n1 = 256, P = 8) with the tuple cut to 2 complex."""
    import torch
    import accfft_amd as af
    ln, P, t = 256, 8, 4
    a = rng(11).standard_normal(ln * P * ln * t).astype(np.float32)
    ad = torch.tensor(a, device=dev)
    packed = torch.empty_like(ad)
    af.local_transpose_col(ln, P, ln * t, ln * t, ad, packed)
    back = torch.empty_like(ad)
    af.local_transpose(P, ln, ln * t, packed, back)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(packed.cpu().numpy(), orc.local_transpose_col(ln, P, ln * t, ln * t, a))
    np.testing.assert_array_equal(back.cpu().numpy(), a)
